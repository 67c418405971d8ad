// CPU oracle for the billboard voxelize -> 3D mip -> voxel cone trace path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference's shaders and
// host code as scalar C++ so that the HIP implementation can be checked
// against it.  Only tests, the smoke check and the benchmark's CPU-baseline
// leg load it; the product library never links or calls it.
//
// Every function cites the reference file:line it follows (paths are relative
// to the reference repository).  The arithmetic is written in a fixed order
// and must be compiled with -ffp-contract=off, so the FP32 results are
// bit-for-bit reproducible by any IEEE-754 implementation that performs the
// same operations in the same order (the GPU kernels do).
//
// Precision model
//   * FP32 pass ("decide"): the exact op sequence the GPU executes.  It takes
//     every discrete decision: fragment coverage (discard), depth-test
//     winners, voxel indices, cone termination step, pixel saturation.
//   * FP64 pass ("replay"): the continuous arithmetic recomputed in double
//     with the FP32 pass's decisions replayed (same volume occupancy, same
//     fragment set, same number of cone samples, same number of composited
//     fragments).  This is the accuracy reference for radiance/alpha.
//
// Parity status of each function is listed in the header comment of the
// function ("pinned by" = the independent check in tests/ that would fail on
// a dropped term, wrong sign/index or transposed operand).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---- plain-C parameter blocks (ctypes mirrors live in tests/oracle_lib.py) ----

typedef struct {
    int dim;            // voxels per side (CloudVolume::dimension)
    int levels;         // mip levels (CloudVolume::levels)
    int format;         // 0 = R8 (reference), 1 = RGBA32F, 2 = RGBA16F
    int mode;           // 0 = reference (shell only, R channel), 1 = rgba (chord fill + shell)
    float bounds_lo[3]; // (xBounds.x, yBounds.x, zBounds.x), relative to position
    float bounds_hi[3]; // (xBounds.y, yBounds.y, zBounds.y)
    float position[3];  // CloudVolume::position
    int noise_octaves;  // 0 = no density modulation
    uint32_t noise_seed;
    float noise_base_freq;
    float noise_lacunarity;
    float noise_gain;
} cvo_volume;

typedef struct {
    float position[3];  // Sun::position (used for the light view, src/Sun.hpp:26-43)
    float direction[3]; // unit vector toward the sun (directional cone direction)
    int directional;    // 0: cone toward the point light at position (reference)
} cvo_light;

typedef struct {
    float eye[3];
    float target[3];
    float up[3];
    float fovy;         // radians (glm::perspective: only tan(fovy / 2) is used)
    float znear;
    float zfar;
    float aspect;       // 0: W / H
    int use_matrices;   // 1: P and V (column-major) replace target/up/fovy/aspect/znear/zfar
    float P[16];
    float V[16];
} cvo_camera;

typedef struct {
    int steps;            // vctSteps
    float cone_angle;     // vctConeAngle (full aperture, radians)
    float initial_height; // vctConeInitialHeight (voxel units)
    float lod_offset;     // vctLodOffset
    float down_scaling;   // vctDownScaling
    int cone_mode;        // 0 weighted R-sum (reference), 1 front-to-back, 2 front-to-back with step extinction
    float alpha_stop;     // early-out threshold (front-to-back modes)
    float sigma;          // extinction per unit texture length (mode 2)
    int composite;        // 0 GL back-to-front blend over all fragments, 1 front-to-back with early-out
    int quantize_rgba8;   // GL blend into an RGBA8 framebuffer (composite 0 only)
    int compensated;      // Kahan-compensated front-to-back accumulation
    float clear_color[4]; // framebuffer clear (composite 0)
} cvo_cone;

// soft-particle noise march (res/conetrace_frag.glsl:29-41 uniforms;
// defaults src/Shaders/ConeTraceShader.hpp:14-25)
typedef struct {
    int enabled;          // doNoiseSample
    uint32_t seed;        // noise-map alpha bytes (a seeded stand-in for rand())
    float step_size;      // stepSize
    float opacity;        // noiseOpacity
    int octaves;          // numOctaves
    float freq_step;      // freqStep
    float pers_step;      // persStep
    float adjust_size;    // adjustSize
    int min_steps;        // minNoiseSteps
    int max_steps;        // maxNoiseSteps
    float min_color;      // minNoiseColor
    float color_scale;    // noiseColorScale
    float offsets[3];     // the octaveOffsets uniform (windVel * runTime)
} cvo_noise_march;

// sun sprite (res/sun_frag.glsl uniforms, defaults src/main.cpp:40-43)
typedef struct {
    int enabled;
    float inner_color[3];
    float outer_color[3];
    float inner_radius;
    float outer_radius;
} cvo_sun_sprite;

typedef struct {
    float V[16];          // glm column-major view matrix
    float near_plane[3];  // Sun::nearPlane
    float clip;           // Sun::clipDistance
    float ortho[4];       // left, right, bottom, top
    float znear, zfar;
} cvo_lightview;

typedef struct {
    int64_t fragments;      // covered (not discarded) camera fragments that were shaded
    int64_t samples;        // cone samples accumulated (FP32 pass)
    int64_t two_level;      // samples that blended two mip levels
    int64_t light_fragments;// covered light-view fragments (pass 1)
    int64_t chord_stores;   // chord-fill voxel stores (pass 1, rgba mode)
} cvo_stats;

}  // extern "C"

namespace {

// ------------------------------------------------------------------------
// small vector helpers (glm 0.9.8 semantics, float)
// ------------------------------------------------------------------------
struct V3 { float x, y, z; };

inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
// glm compute_dot<vec3>: tmp = a*b; return tmp.x + tmp.y + tmp.z
inline float dot(V3 a, V3 b) { float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z; return (x + y) + z; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glm::distance(p0, p1) = length(p1 - p0)
inline float distance(V3 p0, V3 p1) { return length(sub(p1, p0)); }
// glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x)
inline V3 normalize(V3 a) { float inv = 1.0f / std::sqrt(dot(a, a)); return mul(a, inv); }
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

// glm::lookAtRH (column-major, M[col][row] stored at M[col*4+row])
void look_at(V3 eye, V3 center, V3 up, float M[16]) {
    V3 f = normalize(sub(center, eye));
    V3 s = normalize(cross(f, up));
    V3 u = cross(s, f);
    for (int i = 0; i < 16; ++i) M[i] = 0.0f;
    M[0 * 4 + 0] = s.x; M[1 * 4 + 0] = s.y; M[2 * 4 + 0] = s.z;
    M[0 * 4 + 1] = u.x; M[1 * 4 + 1] = u.y; M[2 * 4 + 1] = u.z;
    M[0 * 4 + 2] = -f.x; M[1 * 4 + 2] = -f.y; M[2 * 4 + 2] = -f.z;
    M[3 * 4 + 0] = -dot(s, eye);
    M[3 * 4 + 1] = -dot(u, eye);
    M[3 * 4 + 2] = dot(f, eye);
    M[3 * 4 + 3] = 1.0f;
}

// V * (c, 1): glm mat4*vec4 = ((col0*x + col1*y) + col2*z) + col3*w, w = 1
inline V3 view_point(const float M[16], V3 c) {
    float x = ((M[0] * c.x + M[4] * c.y) + M[8] * c.z) + M[12];
    float y = ((M[1] * c.x + M[5] * c.y) + M[9] * c.z) + M[13];
    float z = ((M[2] * c.x + M[6] * c.y) + M[10] * c.z) + M[14];
    return v3(x, y, z);
}

// Camera-facing billboard basis.  Vi = transpose(V with translation zeroed)
// (src/Shaders/VoxelizeShader.cpp:50-52, src/Shaders/ConeTraceShader.cpp:66-68);
// quad corner = c + s * Vi * (+-1, +-1, 0)  (res/billboard_vert_instanced.glsl:21-31),
// so the in-plane axes are Vi's first two columns = V's first two rows.
// normalize(fragNor) = normalize(Vi * z) (res/billboard_vert_instanced.glsl:34,
// res/first_voxelize.glsl:51) is V's third row, normalised.
struct Basis { V3 S, U, B; };
Basis billboard_basis(const float V[16]) {
    Basis b;
    b.S = v3(V[0], V[4], V[8]);
    b.U = v3(V[1], V[5], V[9]);
    b.B = normalize(v3(V[2], V[6], V[10]));
    return b;
}

// ------------------------------------------------------------------------
// World bounds and the voxel index map
// ------------------------------------------------------------------------
struct Grid {
    int N;
    float lo[3];     // world bounds = position + bounds (src/Shaders/VoxelizeShader.cpp:124-126)
    float range[3];  // xBounds.y - xBounds.x of the world bounds (res/first_voxelize.glsl:24-26)
    float stepSize;  // min(voxelSize) (src/Shaders/VoxelizeShader.cpp:128)
};

Grid make_grid(const cvo_volume* v) {
    Grid g;
    g.N = v->dim;
    float vs[3];
    for (int a = 0; a < 3; ++a) {
        float wlo = v->position[a] + v->bounds_lo[a];
        float whi = v->position[a] + v->bounds_hi[a];
        g.lo[a] = wlo;
        g.range[a] = whi - wlo;
        // CloudVolume::voxelSize = range / (float)dimension with range from the
        // relative bounds (src/CloudVolume.cpp:50-54)
        vs[a] = (v->bounds_hi[a] - v->bounds_lo[a]) / (float)v->dim;
    }
    g.stepSize = std::min(vs[0], std::min(vs[1], vs[2]));
    return g;
}

// calculateVoxelLerp (res/first_voxelize.glsl:23-33):
//   x = voxelDim * ((pos.x - xBounds.x) / rangeX)
inline V3 voxel_lerp(const Grid& g, V3 p) {
    float fN = (float)g.N;
    return v3(fN * ((p.x - g.lo[0]) / g.range[0]),
              fN * ((p.y - g.lo[1]) / g.range[1]),
              fN * ((p.z - g.lo[2]) / g.range[2]));
}

// ivec3(calculateVoxelLerp(pos)) with GL image-store semantics: conversion
// truncates toward zero (so (-1,0) maps to 0) and stores outside [0,N-1] are
// dropped (res/first_voxelize.glsl:35-37, res/second_voxelize.glsl:30-32).
// Returns the x-fastest linear index or -1.
// Pinned by: test_oracle_index_map (closed-form points).
inline int64_t voxel_index(const Grid& g, V3 p) {
    V3 l = voxel_lerp(g, p);
    float fN = (float)g.N;
    if (!(l.x > -1.0f && l.x < fN && l.y > -1.0f && l.y < fN && l.z > -1.0f && l.z < fN)) return -1;
    int64_t ix = (int)l.x, iy = (int)l.y, iz = (int)l.z;
    return ix + (int64_t)g.N * (iy + (int64_t)g.N * iz);
}

// ------------------------------------------------------------------------
// Pixel -> view-plane mapping shared by the light (ortho) and camera
// (perspective) rasterizations.  Billboards are view-aligned, so each lies in
// a plane of constant view depth; a pixel's fragment is the pixel ray's hit
// with that plane (the perspective-correct interpolation of fragPos,
// res/billboard_vert_instanced.glsl:32, evaluated exactly).
// ------------------------------------------------------------------------
inline float pixel_ndc(int p, int n) { return (2.0f * ((float)p + 0.5f)) / (float)n - 1.0f; }

// Sphere distribution and discard (res/first_voxelize.glsl:42-48,
// res/conetrace_frag.glsl:177-182): q = distance(center, fragPos) / radius,
// sc = sqrt(max(0, 1 - q*q)), discard when sc < 0.01.  distance() is the
// in-plane distance sqrt(du*du + dv*dv) of the hit from the centre.
// Pinned by: test_oracle_light_hemisphere (analytic hemisphere under an
// axis-aligned light) and test_oracle_coverage_disc.
inline bool sphere_contrib(float du, float dv, float s, float* sc_out) {
    float dist = std::sqrt(du * du + dv * dv);
    float q = dist / s;
    float sc = std::sqrt(std::max(0.0f, 1.0f - q * q));
    *sc_out = sc;
    return !(sc < 0.01f);
}

struct PixBox { int x0, x1, y0, y1; bool empty; };

// conservative pixel bounding box of the billboard square [c-s, c+s]^2 for a
// view mapping x_view = xn * ex + ox (ex = extent per NDC unit)
PixBox pixel_box(float cx, float cy, float s, float ex, float ey, float ox, float oy, int W, int H) {
    PixBox b;
    float xa = (((cx - s) - ox) / ex + 1.0f) * 0.5f * (float)W - 0.5f;
    float xb = (((cx + s) - ox) / ex + 1.0f) * 0.5f * (float)W - 0.5f;
    float ya = (((cy - s) - oy) / ey + 1.0f) * 0.5f * (float)H - 0.5f;
    float yb = (((cy + s) - oy) / ey + 1.0f) * 0.5f * (float)H - 0.5f;
    float fx0 = std::floor(std::min(xa, xb)) - 1.0f, fx1 = std::ceil(std::max(xa, xb)) + 1.0f;
    float fy0 = std::floor(std::min(ya, yb)) - 1.0f, fy1 = std::ceil(std::max(ya, yb)) + 1.0f;
    b.empty = !(fx1 >= 0.0f && fy1 >= 0.0f && fx0 <= (float)(W - 1) && fy0 <= (float)(H - 1));
    if (b.empty) { b.x0 = b.x1 = b.y0 = b.y1 = 0; return b; }
    b.x0 = (int)std::max(fx0, 0.0f); b.x1 = (int)std::min(fx1, (float)(W - 1));
    b.y0 = (int)std::max(fy0, 0.0f); b.y1 = (int)std::min(fy1, (float)(H - 1));
    return b;
}

// ------------------------------------------------------------------------
// half precision (IEEE binary16) conversions, round-to-nearest-even
// Pinned by: test_oracle_half_conversion (numpy float16 over edge cases).
// ------------------------------------------------------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

uint16_t f2h(float f) {
    uint32_t x = fbits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);          // NaN
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);         // rounds to inf
    if (ax < 0x38800000u) {                                           // half subnormal or zero
        if (ax < 0x33000000u) return (uint16_t)sign;                  // < 2^-25 -> 0 (2^-25 ties to even 0)
        uint32_t e = ax >> 23;                                        // 102..112
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;                                    // 14..24
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t e = (ax >> 23) - 112u;                                   // rebias 127 -> 15
    uint32_t m = ax & 0x7fffffu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;            // may carry into exponent (ok)
    return (uint16_t)(sign | q);
}

float h2f(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0) {
        if (m == 0) return bitsf(sign);
        float v = (float)m * 5.9604644775390625e-8f;                  // m * 2^-24, exact
        return sign ? -v : v;
    }
    if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
    return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

// ------------------------------------------------------------------------
// exp2 for x <= 0 used by the extinction model: 2^x = 2^n * e^(f ln 2),
// n = round-to-nearest of x, |f| <= 1/2, Taylor series of e^z to degree 6
// (|z| <= 0.347, truncation < 2e-7 relative).  Same op sequence in the
// HIP kernels so the FP32 paths agree bit for bit; the FP64 pass uses
// std::exp2.  Pinned by: test_oracle_exp2 (against math.exp2).
// ------------------------------------------------------------------------
inline float cv_exp2_neg(float x) {
    if (!(x > -126.0f)) return 0.0f;
    float n = std::floor(x + 0.5f);
    float f = x - n;
    float z = f * 0.693147180559945309f;
    float p = 1.0f / 720.0f;  // Horner with fused multiply-adds
    p = std::fma(p, z, 1.0f / 120.0f);
    p = std::fma(p, z, 1.0f / 24.0f);
    p = std::fma(p, z, 1.0f / 6.0f);
    p = std::fma(p, z, 0.5f);
    p = std::fma(p, z, 1.0f);
    p = std::fma(p, z, 1.0f);
    int ni = (int)n;
    float scale = bitsf((uint32_t)(ni + 127) << 23);
    return p * scale;
}

// ------------------------------------------------------------------------
// value noise (density modulation of the precision variant)
// lattice value = murmur3 fmix32 of a Teschner-style spatial hash
// ------------------------------------------------------------------------
inline uint32_t lattice_hash(uint32_t seed, int ix, int iy, int iz) {
    uint32_t h = seed;
    h ^= (uint32_t)ix * 0x8da6b343u;
    h ^= (uint32_t)iy * 0xd8163841u;
    h ^= (uint32_t)iz * 0xcb1ab31fu;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
template <class T> inline T lattice_value(uint32_t seed, int ix, int iy, int iz) {
    return (T)(lattice_hash(seed, ix, iy, iz) >> 8) * (T)(1.0 / 16777216.0);
}
template <class T> inline T smooth(T f) { return (f * f) * ((T)3 - (T)2 * f); }
template <class T> inline T lerpT(T a, T b, T t) { return a * ((T)1 - t) + b * t; }

// one octave: trilinear interpolation of lattice values with smoothstep weights
template <class T> T value_noise(uint32_t seed, T px, T py, T pz) {
    T fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    T wx = smooth(px - fx), wy = smooth(py - fy), wz = smooth(pz - fz);
    T c000 = lattice_value<T>(seed, ix, iy, iz), c100 = lattice_value<T>(seed, ix + 1, iy, iz);
    T c010 = lattice_value<T>(seed, ix, iy + 1, iz), c110 = lattice_value<T>(seed, ix + 1, iy + 1, iz);
    T c001 = lattice_value<T>(seed, ix, iy, iz + 1), c101 = lattice_value<T>(seed, ix + 1, iy, iz + 1);
    T c011 = lattice_value<T>(seed, ix, iy + 1, iz + 1), c111 = lattice_value<T>(seed, ix + 1, iy + 1, iz + 1);
    T x00 = lerpT(c000, c100, wx), x10 = lerpT(c010, c110, wx);
    T x01 = lerpT(c001, c101, wx), x11 = lerpT(c011, c111, wx);
    T y0 = lerpT(x00, x10, wy), y1 = lerpT(x01, x11, wy);
    return lerpT(y0, y1, wz);
}

// fractal sum at voxel (x,y,z) of an N^3 grid, normalised to [0,1)
template <class T> T fbm_voxel(const cvo_volume* v, int x, int y, int z) {
    T fN = (T)v->dim;
    T u = ((T)x + (T)0.5) / fN, w = ((T)y + (T)0.5) / fN, t = ((T)z + (T)0.5) / fN;
    T freq = (T)v->noise_base_freq, amp = (T)1, sum = (T)0, norm = (T)0;
    for (int o = 0; o < v->noise_octaves; ++o) {
        uint32_t seed = v->noise_seed + (uint32_t)o * 0x9e3779b9u;
        sum = sum + amp * value_noise<T>(seed, u * freq, w * freq, t * freq);
        norm = norm + amp;
        freq = freq * (T)v->noise_lacunarity;
        amp = amp * (T)v->noise_gain;
    }
    return sum / norm;
}

// ------------------------------------------------------------------------
// Level storage (linear, x-fastest), in native format and in double
// ------------------------------------------------------------------------
inline int level_dim(int N, int k) { int d = N >> k; return d < 1 ? 1 : d; }
inline int channels(int format) { return format == 0 ? 1 : 4; }
inline size_t elem_size(int format) { return format == 0 ? 1 : (format == 1 ? 4 : 2); }

std::vector<size_t> level_offsets(int N, int L, int C) {   // in elements
    std::vector<size_t> off(L + 1, 0);
    for (int k = 0; k < L; ++k) {
        size_t d = (size_t)level_dim(N, k);
        off[k + 1] = off[k] + d * d * d * (size_t)C;
    }
    return off;
}

// parent texel p of a level with child size n covers children [2p, 2p+1],
// except that the last parent of an odd-sized child level also takes the
// final child (nothing is dropped for non-power-of-two sizes).
inline void child_range(int p, int nparent, int nchild, int* c0, int* c1) {
    *c0 = std::min(2 * p, nchild - 1);
    *c1 = std::min(2 * p + 1, nchild - 1);
    if (p == nparent - 1) *c1 = nchild - 1;
}

}  // namespace

// ========================================================================
// exported API
// ========================================================================
extern "C" {

// Sun::update (src/Sun.hpp:26-43): light view from the sun toward the volume.
// The up vector is (0,1,0) as in the reference; when the look direction is
// within 0.999 of +-y (where lookAt degenerates) (0,0,1) is used instead.
// Pinned by: test_oracle_sun_update (closed form for an axis-aligned sun).
void cvo_sun_update(const cvo_volume* vol, const float sun_pos[3], cvo_lightview* out) {
    V3 mn = v3(vol->bounds_lo[0], vol->bounds_lo[1], vol->bounds_lo[2]);
    V3 mx = v3(vol->bounds_hi[0], vol->bounds_hi[1], vol->bounds_hi[2]);
    V3 vpos = v3(vol->position[0], vol->position[1], vol->position[2]);
    V3 spos = v3(sun_pos[0], sun_pos[1], sun_pos[2]);
    V3 lookDir = normalize(sub(vpos, spos));
    float minLen = length(mn), maxLen = length(mx);
    float ml = std::max(minLen, maxLen);
    V3 lookPos = sub(vpos, mul(lookDir, ml));
    V3 up = v3(0.0f, 1.0f, 0.0f);
    if (std::fabs(lookDir.y) > 0.999f) up = v3(0.0f, 0.0f, 1.0f);
    look_at(lookPos, vpos, up, out->V);
    float minmin = 2.0f * std::min(mn.x, std::min(mn.y, mn.z));
    float maxmax = 2.0f * std::max(mx.x, std::max(mx.y, mx.z));
    V3 nearPlane = add(lookPos, mul(lookDir, 0.01f));
    V3 farPlane = add(lookPos, mul(mul(lookDir, 2.0f), ml));
    float clip = distance(nearPlane, farPlane);
    out->near_plane[0] = nearPlane.x; out->near_plane[1] = nearPlane.y; out->near_plane[2] = nearPlane.z;
    out->clip = clip;
    out->ortho[0] = minmin; out->ortho[1] = maxmax; out->ortho[2] = minmin; out->ortho[3] = maxmax;
    out->znear = 0.01f;
    out->zfar = 0.01f + clip;
}

// Both voxelization passes (src/Shaders/VoxelizeShader.cpp:18-118):
//   pass 1 (res/first_voxelize.glsl:39-64) over billboards in upload order:
//     sphere point nearest the light, depth = distance(lightNearPlane, p) / clip,
//     depth test LESS against 1.0, first drawn wins ties; rgba mode also runs the
//     chord fill of res/first_voxelize.glsl:53-58 (commented out in the shipped
//     shader; paper/tex/voxelization.tex:12-27) for every covered fragment;
//   pass 2 (res/second_voxelize.glsl:34-51): each texel with a > 0 stores 1 at
//     its voxel and at the 8 diagonal neighbours stepSize * normalize(+-1,+-1,+-1).
// shell/fill: N^3 bytes (x-fastest), set to 0/1.  posmap: W*H*4 floats or NULL.
// winner: W*H ints (sphere index, -1) or NULL.  radius[] is scale * fluffiness.
// Pinned by: test_oracle_light_hemisphere, test_oracle_depth_ties,
// test_oracle_splat_bruteforce, test_oracle_chord_bruteforce.
int cvo_voxelize(const cvo_volume* vol, int S, const float* xyz, const float* radius,
                 const cvo_light* light, int W, int H, uint8_t* shell, uint8_t* fill,
                 float* posmap, int32_t* winner, cvo_stats* stats) {
    if (!vol || vol->dim <= 0 || W <= 0 || H <= 0 || S < 0) return -1;
    Grid g = make_grid(vol);
    size_t nvox = (size_t)g.N * g.N * g.N;
    std::memset(shell, 0, nvox);
    if (fill) std::memset(fill, 0, nvox);
    cvo_lightview lv;
    cvo_sun_update(vol, light->position, &lv);
    Basis bb = billboard_basis(lv.V);
    V3 np = v3(lv.near_plane[0], lv.near_plane[1], lv.near_plane[2]);
    float lsx = (lv.ortho[1] - lv.ortho[0]) * 0.5f, lox = (lv.ortho[1] + lv.ortho[0]) * 0.5f;
    float lsy = (lv.ortho[3] - lv.ortho[2]) * 0.5f, loy = (lv.ortho[3] + lv.ortho[2]) * 0.5f;

    std::vector<float> depth((size_t)W * H, 1.0f);
    std::vector<int32_t> win((size_t)W * H, -1);
    std::vector<float> wpx((size_t)W * H * 3, 0.0f);
    int64_t nlight = 0, nchord = 0;
    const bool chord = (vol->mode == 1) && fill;

    // rows are independent: each thread owns a band of rows and visits every
    // billboard in draw order, so ties resolve exactly as in GL
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nlight, nchord)
    for (int py = 0; py < H; ++py) {
        float yn = pixel_ndc(py, H);
        float hy = yn * lsy + loy;
        for (int j = 0; j < S; ++j) {
            V3 c = v3(vol->position[0] + xyz[3 * j], vol->position[1] + xyz[3 * j + 1],
                      vol->position[2] + xyz[3 * j + 2]);
            float s = radius[j];
            if (!(s > 0.0f)) continue;
            V3 cv = view_point(lv.V, c);
            float dz = -cv.z;
            if (!(dz >= lv.znear && dz <= lv.zfar)) continue;
            PixBox box = pixel_box(cv.x, cv.y, s, lsx, lsy, lox, loy, W, H);
            if (box.empty || py < box.y0 || py > box.y1) continue;
            float dv = hy - cv.y;
            for (int px = box.x0; px <= box.x1; ++px) {
                float xn = pixel_ndc(px, W);
                float hx = xn * lsx + lox;
                float du = hx - cv.x;
                float sc;
                if (!sphere_contrib(du, dv, s, &sc)) continue;
                ++nlight;
                // fragPos on the billboard plane
                V3 fp = v3((c.x + du * bb.S.x) + dv * bb.U.x, (c.y + du * bb.S.y) + dv * bb.U.y,
                           (c.z + du * bb.S.z) + dv * bb.U.z);
                float dist = s * sc;  // radius * sphereContrib (res/first_voxelize.glsl:52)
                if (chord) {
                    // start = fragPos - dir * dist; for (i = 0; i < 2*dist; i += stepSize)
                    V3 start = sub(fp, mul(bb.B, dist));
                    float lim = 2.0f * dist;
                    for (float i = 0.0f; i < lim; i += g.stepSize) {
                        V3 p = add(start, mul(bb.B, i));
                        int64_t idx = voxel_index(g, p);
                        ++nchord;
                        if (idx >= 0) __atomic_store_n(&fill[idx], (uint8_t)1, __ATOMIC_RELAXED);
                    }
                }
                V3 wp = add(fp, mul(bb.B, dist));  // fragPos + dir * dist (:61)
                float d = distance(np, wp) / lv.clip;  // (:63)
                d = std::min(std::max(d, 0.0f), 1.0f);
                size_t pix = (size_t)py * W + px;
                if (d < depth[pix]) {  // GL_LESS, cleared to 1.0; ties keep the first drawn
                    depth[pix] = d;
                    win[pix] = j;
                    wpx[3 * pix] = wp.x; wpx[3 * pix + 1] = wp.y; wpx[3 * pix + 2] = wp.z;
                }
            }
        }
    }

    // pass 2 (res/second_voxelize.glsl:34-51); normalize(vec3(1,1,1)) = 1/sqrt(3) in FP32
    const float k = 1.0f / std::sqrt(3.0f);
    const float o = g.stepSize * k;
    static const int sgn[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                                  {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            size_t pix = (size_t)py * W + px;
            if (posmap) {
                posmap[4 * pix] = wpx[3 * pix]; posmap[4 * pix + 1] = wpx[3 * pix + 1];
                posmap[4 * pix + 2] = wpx[3 * pix + 2]; posmap[4 * pix + 3] = win[pix] >= 0 ? 1.0f : 0.0f;
            }
            if (winner) winner[pix] = win[pix];
            if (win[pix] < 0) continue;
            V3 wp = v3(wpx[3 * pix], wpx[3 * pix + 1], wpx[3 * pix + 2]);
            int64_t idx = voxel_index(g, wp);
            if (idx >= 0) __atomic_store_n(&shell[idx], (uint8_t)1, __ATOMIC_RELAXED);
            for (int t = 0; t < 8; ++t) {
                V3 off = v3(sgn[t][0] > 0 ? o : -o, sgn[t][1] > 0 ? o : -o, sgn[t][2] > 0 ? o : -o);
                idx = voxel_index(g, add(wp, off));
                if (idx >= 0) __atomic_store_n(&shell[idx], (uint8_t)1, __ATOMIC_RELAXED);
            }
        }
    }
    if (stats) { stats->light_fragments = nlight; stats->chord_stores = nchord; }
    return 0;
}

// Occupancy count and order-independent hash over level 0.  state 2 = lit
// shell (pass 2), 1 = chord fill only, 0 = empty.  hash = sum over occupied
// voxels of splitmix64(index * 4 + state) mod 2^64.  The reference's own
// occupancy figure is VoxelShader's activeVoxels count (src/Shaders/VoxelShader.cpp:110-115).
void cvo_occupancy(int N, const uint8_t* shell, const uint8_t* fill, uint64_t* count, uint64_t* hash) {
    uint64_t c = 0, h = 0;
    size_t nvox = (size_t)N * N * N;
    for (size_t i = 0; i < nvox; ++i) {
        uint64_t st = shell[i] ? 2 : ((fill && fill[i]) ? 1 : 0);
        if (!st) continue;
        ++c;
        uint64_t z = (uint64_t)i * 4u + st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        h += z;
    }
    *count = c;
    *hash = h;
}

// Level 0 from occupancy, then the mip chain (glGenerateMipmap,
// src/Shaders/VoxelizeShader.cpp:105; storage src/CloudVolume.cpp:18-23).
//   reference mode: R = 1 on the lit shell (image stores of vec4(1) into GL_R8)
//   rgba mode: shell = (1,1,1,1), chord fill only = (0,0,0,1)  (paper semantics)
//   noise: A *= fbm(voxel centre) when noise_octaves > 0 (rgba mode)
// mips: parent = sum of children in z, y, x order (x fastest) / child count,
//   R8 parents = (sum of 8-bit values + count/2) / count (integer, round half up).
// out_native: all levels concatenated, linear x-fastest, element type by
// format (u8 / float / binary16).  out_f64 (optional): exact mean in double.
// Pinned by: test_oracle_mips_constant, test_oracle_mips_single_voxel,
// test_oracle_mips_mean (mean preserved for power-of-two N).
int cvo_build_levels(const cvo_volume* vol, const uint8_t* shell, const uint8_t* fill,
                     void* out_native, double* out_f64) {
    int N = vol->dim, L = vol->levels, C = channels(vol->format);
    if (N <= 0 || L <= 0) return -1;
    std::vector<size_t> off = level_offsets(N, L, C);
    std::vector<float> lf(off[L]);  // float values (decoded native)
    const bool noise = vol->mode == 1 && vol->noise_octaves > 0;
    // level 0
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = (size_t)x + (size_t)N * (y + (size_t)N * z);
                float r, gch, b, a;
                double rd, gd, bd, ad;
                if (vol->mode == 0) {
                    float v = shell[i] ? 1.0f : 0.0f;
                    r = gch = b = a = v;
                    rd = gd = bd = ad = v;
                } else {
                    if (shell[i]) { r = gch = b = a = 1.0f; }
                    else if (fill && fill[i]) { r = gch = b = 0.0f; a = 1.0f; }
                    else { r = gch = b = a = 0.0f; }
                    rd = r; gd = gch; bd = b; ad = a;
                    if (noise && a != 0.0f) {
                        a = a * fbm_voxel<float>(vol, x, y, z);
                        ad = ad * fbm_voxel<double>(vol, x, y, z);
                    }
                }
                if (C == 1) {
                    lf[i] = r;
                    if (out_f64) out_f64[i] = rd;
                } else {
                    float vals[4] = {r, gch, b, a};
                    double vd[4] = {rd, gd, bd, ad};
                    for (int ch = 0; ch < 4; ++ch) {
                        float q = vals[ch];
                        if (vol->format == 2) q = h2f(f2h(q));
                        lf[4 * i + ch] = q;
                        if (out_f64) out_f64[4 * i + ch] = vd[ch];
                    }
                }
            }
    // R8 keeps integer texels
    std::vector<uint8_t> l8;
    if (vol->format == 0) {
        l8.resize(off[L]);
        for (size_t i = 0; i < off[1]; ++i) l8[i] = lf[i] != 0.0f ? 255 : 0;
    }
    for (int k = 1; k < L; ++k) {
        int nc = level_dim(N, k - 1), np_ = level_dim(N, k);
        size_t oc = off[k - 1], op = off[k];
#pragma omp parallel for schedule(static)
        for (int z = 0; z < np_; ++z)
            for (int y = 0; y < np_; ++y)
                for (int x = 0; x < np_; ++x) {
                    int x0, x1, y0, y1, z0, z1;
                    child_range(x, np_, nc, &x0, &x1);
                    child_range(y, np_, nc, &y0, &y1);
                    child_range(z, np_, nc, &z0, &z1);
                    int cnt = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
                    size_t pi = (size_t)x + (size_t)np_ * (y + (size_t)np_ * z);
                    for (int ch = 0; ch < C; ++ch) {
                        float sum = 0.0f;
                        double sumd = 0.0;
                        uint32_t sum8 = 0;
                        for (int cz = z0; cz <= z1; ++cz)
                            for (int cy = y0; cy <= y1; ++cy)
                                for (int cx = x0; cx <= x1; ++cx) {
                                    size_t ci = (size_t)cx + (size_t)nc * (cy + (size_t)nc * cz);
                                    sum = sum + lf[oc + C * ci + ch];
                                    if (out_f64) sumd += out_f64[oc + C * ci + ch];
                                    if (vol->format == 0) sum8 += l8[oc + ci];
                                }
                        if (vol->format == 0) {
                            uint32_t v8 = (sum8 + (uint32_t)cnt / 2u) / (uint32_t)cnt;
                            l8[op + pi] = (uint8_t)v8;
                            lf[op + pi] = (float)v8 / 255.0f;
                        } else {
                            float v = sum / (float)cnt;
                            if (vol->format == 2) v = h2f(f2h(v));
                            lf[op + C * pi + ch] = v;
                        }
                        if (out_f64) out_f64[op + C * pi + ch] = sumd / (double)cnt;
                    }
                }
    }
    if (vol->format == 0) {
        std::memcpy(out_native, l8.data(), off[L]);
    } else if (vol->format == 1) {
        std::memcpy(out_native, lf.data(), off[L] * sizeof(float));
    } else {
        uint16_t* o = (uint16_t*)out_native;
        for (size_t i = 0; i < off[L]; ++i) o[i] = f2h(lf[i]);
    }
    return 0;
}

// CloudVolume::sortBoards (src/CloudVolume.cpp:65-82), literally: selection
// sort by descending glm::distance(position + p, point); the first maximum
// wins; elements are swapped (not stable).  order[r] = original index drawn r-th.
// Pinned by: test_oracle_sort (matches a stable descending sort on distinct keys).
void cvo_sort_boards(int S, const float* xyz, const float vol_pos[3], const float point[3], int32_t* order) {
    std::vector<int32_t> idx(S);
    std::vector<float> dist(S);
    V3 pt = v3(point[0], point[1], point[2]);
    for (int j = 0; j < S; ++j) {
        idx[j] = j;
        V3 c = v3(vol_pos[0] + xyz[3 * j], vol_pos[1] + xyz[3 * j + 1], vol_pos[2] + xyz[3 * j + 2]);
        dist[j] = distance(c, pt);
    }
    for (int i = 0; i < S; ++i) {
        int m = i;
        for (int j = i + 1; j < S; ++j)
            if (dist[m] < dist[j]) m = j;
        if (m != i) { std::swap(dist[i], dist[m]); std::swap(idx[i], idx[m]); }
    }
    for (int i = 0; i < S; ++i) order[i] = idx[i];
}

// FP32 distance key used by the sort (for tests)
void cvo_board_distances(int S, const float* xyz, const float vol_pos[3], const float point[3], float* out) {
    V3 pt = v3(point[0], point[1], point[2]);
    for (int j = 0; j < S; ++j) {
        V3 c = v3(vol_pos[0] + xyz[3 * j], vol_pos[1] + xyz[3 * j + 1], vol_pos[2] + xyz[3 * j + 2]);
        out[j] = distance(c, pt);
    }
}

void cvo_f2h_array(const float* in, uint16_t* out, int64_t n) { for (int64_t i = 0; i < n; ++i) out[i] = f2h(in[i]); }
void cvo_h2f_array(const uint16_t* in, float* out, int64_t n) { for (int64_t i = 0; i < n; ++i) out[i] = h2f(in[i]); }
void cvo_exp2_neg_array(const float* in, float* out, int64_t n) { for (int64_t i = 0; i < n; ++i) out[i] = cv_exp2_neg(in[i]); }
void cvo_noise_voxels(const cvo_volume* v, const int32_t* xyz, int64_t n, float* out32, double* out64) {
    for (int64_t i = 0; i < n; ++i) {
        out32[i] = fbm_voxel<float>(v, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        if (out64) out64[i] = fbm_voxel<double>(v, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    }
}
void cvo_voxel_index(const cvo_volume* vol, const float* pts, int64_t n, int64_t* out) {
    Grid g = make_grid(vol);
    for (int64_t i = 0; i < n; ++i) out[i] = voxel_index(g, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}

}  // extern "C"

// ========================================================================
// camera pass: cone trace + composite
// ========================================================================
namespace {

template <class T> struct Vec3T { T x, y, z; };

// one mip level view over a linear array
template <class T> struct Level { int n; size_t off; };

struct Volume32 {           // FP32 view of the native storage
    int N, L, C, format;
    const void* data;
    std::vector<size_t> off;
    inline float texel(int k, int x, int y, int z, int ch) const {
        int n = level_dim(N, k);
        size_t i = off[k] + (size_t)C * ((size_t)x + (size_t)n * (y + (size_t)n * z)) + ch;
        if (format == 0) return (float)((const uint8_t*)data)[i] / 255.0f;
        if (format == 1) return ((const float*)data)[i];
        return h2f(((const uint16_t*)data)[i]);
    }
};
struct Volume64 {
    int N, L, C;
    const double* data;
    std::vector<size_t> off;
    inline double texel(int k, int x, int y, int z, int ch) const {
        int n = level_dim(N, k);
        return data[off[k] + (size_t)C * ((size_t)x + (size_t)n * (y + (size_t)n * z)) + ch];
    }
};

// IEEE maxNum/minNum clamp (a NaN operand yields the other one), as fmaxf/fminf
template <class T> inline T clampT(T v, T lo, T hi) { return std::fmin(std::fmax(v, lo), hi); }

// GL 4.4 8.14.2 trilinear filtering of one level with CLAMP_TO_EDGE:
//   u = s * w - 1/2, i0 = floor(u), alpha = frac(u), i1 = i0 + 1, both clamped;
//   evaluated as lerps along x, then y, then z, each lerp(a, b, t) =
//   fma(t, b - a, a) (GL leaves the filter arithmetic to the implementation;
//   a fused multiply-add is correctly rounded on every IEEE target).
// Coordinates are first clamped to [-1, 2] (outside [0,1] every tap is a
// border texel already; the clamp keeps the arithmetic finite).
// Pinned by: test_oracle_sampler_centres, test_oracle_sampler_affine,
// test_oracle_sampler_clamp.
template <class T> inline T lerpF(T a, T b, T t) { return std::fma(t, b - a, a); }

template <class T, class VOL>
void trilinear(const VOL& vol, int k, T sx, T sy, T sz, int nch, const int* chs, T* out) {
    int n = level_dim(vol.N, k);
    T fn = (T)n;
    T u[3] = {std::fma(clampT(sx, (T)-1, (T)2), fn, (T)-0.5), std::fma(clampT(sy, (T)-1, (T)2), fn, (T)-0.5),
              std::fma(clampT(sz, (T)-1, (T)2), fn, (T)-0.5)};
    int i0[3], i1[3];
    T a[3];
    for (int ax = 0; ax < 3; ++ax) {
        T fl = std::floor(u[ax]);
        a[ax] = u[ax] - fl;
        int i = (int)fl;
        i0[ax] = std::min(std::max(i, 0), n - 1);
        i1[ax] = std::min(std::max(i + 1, 0), n - 1);
    }
    for (int q = 0; q < nch; ++q) {
        int ch = chs[q];
        T t000 = (T)vol.texel(k, i0[0], i0[1], i0[2], ch), t100 = (T)vol.texel(k, i1[0], i0[1], i0[2], ch);
        T t010 = (T)vol.texel(k, i0[0], i1[1], i0[2], ch), t110 = (T)vol.texel(k, i1[0], i1[1], i0[2], ch);
        T t001 = (T)vol.texel(k, i0[0], i0[1], i1[2], ch), t101 = (T)vol.texel(k, i1[0], i0[1], i1[2], ch);
        T t011 = (T)vol.texel(k, i0[0], i1[1], i1[2], ch), t111 = (T)vol.texel(k, i1[0], i1[1], i1[2], ch);
        T c00 = lerpF(t000, t100, a[0]);
        T c10 = lerpF(t010, t110, a[0]);
        T c01 = lerpF(t001, t101, a[0]);
        T c11 = lerpF(t011, t111, a[0]);
        T c0 = lerpF(c00, c10, a[1]);
        T c1 = lerpF(c01, c11, a[1]);
        out[q] = lerpF(c0, c1, a[2]);
    }
}

// per-step schedule of traceCone (res/conetrace_frag.glsl:70-76):
//   coneRadius = coneHeight * tan(coneAngle / 2); lod = log2(max(1, 2 * coneRadius))
//   coneHeight += coneRadius
// It depends only on the step index, so it is tabulated once per call.
// lambda -> (level, frac) per GL 4.4 8.14.3 for LINEAR_MIPMAP_LINEAR with the
// magnification switch at lambda <= 0 and clamping at the last level.
template <class T> struct StepTab {
    std::vector<T> h, r, kk;
    std::vector<int> lvl;
    std::vector<T> frac;
    T tanHalf;
};

template <class T> StepTab<T> make_steps(const cvo_cone* cp, int L, int N) {
    StepTab<T> t;
    T angle = (T)cp->cone_angle;
    T tanHalf = std::tan(angle / (T)2);
    t.tanHalf = tanHalf;
    T h = (T)cp->initial_height;
    // extinction per voxel step in log2 units: sigma * log2(e) / N (texture length)
    T kscale = ((T)cp->sigma * (T)1.4426950408889634) / (T)N;
    for (int i = 1; i <= cp->steps; ++i) {
        T r = h * tanHalf;
        T lod = std::log2(std::max((T)1, (T)2 * r));
        T lam = lod + (T)cp->lod_offset;
        int lv;
        T f;
        if (lam <= (T)0) { lv = 0; f = (T)0; }
        else if (lam >= (T)(L - 1)) { lv = L - 1; f = (T)0; }
        else { T fl = std::floor(lam); lv = (int)fl; f = lam - fl; }
        t.h.push_back(h); t.r.push_back(r); t.lvl.push_back(lv); t.frac.push_back(f);
        t.kk.push_back(r * kscale);
        h = h + r;
    }
    return t;
}

// textureLod(volume, p, lod) (GL 4.4 8.14-8.15): (1 - f) * tau1 + f * tau2,
// evaluated as lerp(tau1, tau2, f)
template <class T, class VOL>
void sample_lod(const VOL& vol, const StepTab<T>& st, int i, T sx, T sy, T sz, int nch, const int* chs, T* out) {
    int lv = st.lvl[i];
    T f = st.frac[i];
    trilinear<T>(vol, lv, sx, sy, sz, nch, chs, out);
    if (f != (T)0) {
        T o2[4];
        trilinear<T>(vol, lv + 1, sx, sy, sz, nch, chs, o2);
        for (int q = 0; q < nch; ++q) out[q] = lerpF(out[q], o2[q], f);
    }
}

template <class T> inline T exp2_neg(T x);
template <> inline float exp2_neg<float>(float x) { return cv_exp2_neg(x); }
template <> inline double exp2_neg<double>(double x) { return std::exp2(x); }

template <class T> struct ConeResult { T rgba[4]; int nsamp; };

// traceCone (res/conetrace_frag.glsl:64-79).
//   mode 0 (reference, R11): color += s.r * float(i) / (steps * downScaling)
//     over all steps, no exit test; fragment = vec4(color).
//   modes 1/2 (front-to-back): stop before a sample outside [0,1]^3, accumulate
//     w = (1 - a) * alpha_i;  c += w * s.rgb;  a += w;  stop once a >= alpha_stop.
//     mode 1: alpha_i = s.a;  mode 2: alpha_i = 1 - 2^(-kk_i * s.a) with
//     kk_i = sigma * log2(e) * coneRadius_i / N (Beer-Lambert over the step).
//     compensated: Kahan summation of c and a.
// replay < 0: decide the number of samples (FP32); else accumulate exactly
// `replay` samples (FP64 replay of the FP32 decision).
// Pinned by: test_oracle_cone_constant_weighted (v*(steps+1)/(2*ds)),
// test_oracle_cone_ftb_constant (1-(1-alpha)^n and the early-out step),
// test_oracle_cone_empty, test_oracle_cone_height.
template <class T, class VOL>
ConeResult<T> trace_cone(const VOL& vol, const StepTab<T>& st, const cvo_cone* cp, Vec3T<T> p, Vec3T<T> d,
                         int replay) {
    ConeResult<T> res;
    res.nsamp = 0;
    for (int q = 0; q < 4; ++q) res.rgba[q] = (T)0;
    if (cp->cone_mode == 0) {
        static const int chR[1] = {0};
        T color = (T)0;
        T den = (T)cp->steps * (T)cp->down_scaling;
        int n = replay >= 0 ? replay : cp->steps;
        for (int i = 1; i <= n; ++i) {
            T h = st.h[i - 1];
            T s;
            sample_lod<T>(vol, st, i - 1, std::fma(h, d.x, p.x), std::fma(h, d.y, p.y), std::fma(h, d.z, p.z), 1, chR, &s);
            color = color + (s * (T)i) / den;
        }
        res.nsamp = n;
        for (int q = 0; q < 4; ++q) res.rgba[q] = color;
        return res;
    }
    static const int chRGBA[4] = {0, 1, 2, 3};
    T c[3] = {(T)0, (T)0, (T)0}, a = (T)0;
    T cc[3] = {(T)0, (T)0, (T)0}, ca = (T)0;  // Kahan compensation terms
    T stop = (T)cp->alpha_stop;
    int n = 0;
    for (int i = 1; i <= cp->steps; ++i) {
        if (replay >= 0 && n >= replay) break;
        T h = st.h[i - 1];
        T sx = std::fma(h, d.x, p.x), sy = std::fma(h, d.y, p.y), sz = std::fma(h, d.z, p.z);
        if (replay < 0 && !(sx >= (T)0 && sx <= (T)1 && sy >= (T)0 && sy <= (T)1 && sz >= (T)0 && sz <= (T)1))
            break;
        T s[4];
        sample_lod<T>(vol, st, i - 1, sx, sy, sz, 4, chRGBA, s);
        T al;
        if (cp->cone_mode == 1) al = s[3];
        else al = (T)1 - exp2_neg<T>(-(st.kk[i - 1] * s[3]));
        T w = ((T)1 - a) * al;
        if (cp->compensated) {
            for (int q = 0; q < 3; ++q) {
                T y = std::fma(w, s[q], -cc[q]);
                T t = c[q] + y;
                cc[q] = (t - c[q]) - y;
                c[q] = t;
            }
            T y = w - ca;
            T t = a + y;
            ca = (t - a) - y;
            a = t;
        } else {
            for (int q = 0; q < 3; ++q) c[q] = std::fma(w, s[q], c[q]);
            a = a + w;
        }
        ++n;
        if (replay < 0 && a >= stop) break;
    }
    res.nsamp = n;
    res.rgba[0] = c[0]; res.rgba[1] = c[1]; res.rgba[2] = c[2]; res.rgba[3] = a;
    return res;
}

struct CamSetup {
    float V[16];
    Basis bb;
    float kx, ky;     // tan(fovy/2) * aspect, tan(fovy/2)
    float znear, zfar;
    double Vd[16];
    Vec3T<double> Sd, Ud, Bd;
    double kxd, kyd;
};

// glm::lookAt / perspective in double for the replay
void look_at_d(const float e[3], const float t[3], const float u[3], double M[16], Vec3T<double>* S,
               Vec3T<double>* U, Vec3T<double>* B) {
    double ex = e[0], ey = e[1], ez = e[2];
    double fx = t[0] - ex, fy = t[1] - ey, fz = t[2] - ez;
    double fl = std::sqrt(fx * fx + fy * fy + fz * fz);
    fx /= fl; fy /= fl; fz /= fl;
    double sx = fy * u[2] - u[1] * fz, sy = fz * u[0] - u[2] * fx, sz = fx * u[1] - u[0] * fy;
    double sl = std::sqrt(sx * sx + sy * sy + sz * sz);
    sx /= sl; sy /= sl; sz /= sl;
    double ux = sy * fz - fy * sz, uy = sz * fx - fz * sx, uz = sx * fy - fx * sy;
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = sx; M[4] = sy; M[8] = sz;
    M[1] = ux; M[5] = uy; M[9] = uz;
    M[2] = -fx; M[6] = -fy; M[10] = -fz;
    M[12] = -(sx * ex + sy * ey + sz * ez);
    M[13] = -(ux * ex + uy * ey + uz * ez);
    M[14] = fx * ex + fy * ey + fz * ez;
    M[15] = 1.0;
    S->x = sx; S->y = sy; S->z = sz;
    U->x = ux; U->y = uy; U->z = uz;
    B->x = -fx; B->y = -fy; B->z = -fz;
}

// The camera as ConeTraceShader uploads it (src/Shaders/ConeTraceShader.cpp:64-69,
// src/Camera.cpp:59-60): V = lookAt(eye, target, up) and glm::perspective's
// tanHalfFovy = tan(fovy / 2) (float), the x scale tanHalfFovy * aspect; or V
// and P given as matrices, the scales then 1 / P[0][0] and 1 / P[1][1] and the
// depth range n = P[3][2] / (P[2][2] - 1), f = P[3][2] / (P[2][2] + 1) (the
// inverse of glm::perspective's P[2][2] = -(f + n) / (f - n),
// P[3][2] = -2 f n / (f - n)).  Returns false on an unusable camera.
// Pinned by: test_oracle_camera_hemisphere (start points on the camera-facing
// hemisphere), test_oracle_camera_matrices (matrices equal to lookAt /
// perspective give the same image), test_oracle_reference_projection.
bool camera_setup(const cvo_camera* cam, int W, int H, CamSetup* cs) {
    if (cam->use_matrices) {
        const float* P = cam->P;
        static const int zero[] = {1, 2, 3, 4, 6, 7, 8, 9, 12, 13, 15};
        for (int i : zero)
            if (P[i] != 0.0f) return false;
        if (P[11] != -1.0f || !(P[0] > 0.0f) || !(P[5] > 0.0f)) return false;
        for (int i = 0; i < 16; ++i) { cs->V[i] = cam->V[i]; cs->Vd[i] = (double)cam->V[i]; }
        cs->kx = 1.0f / P[0];
        cs->ky = 1.0f / P[5];
        cs->znear = P[14] / (P[10] - 1.0f);
        cs->zfar = P[14] / (P[10] + 1.0f);
        if (!(cs->znear > 0.0f) || !(cs->zfar > cs->znear)) return false;
        cs->kxd = 1.0 / (double)P[0];
        cs->kyd = 1.0 / (double)P[5];
        const double* M = cs->Vd;
        cs->Sd = Vec3T<double>{M[0], M[4], M[8]};
        cs->Ud = Vec3T<double>{M[1], M[5], M[9]};
        double bl = std::sqrt(M[2] * M[2] + M[6] * M[6] + M[10] * M[10]);
        cs->Bd = Vec3T<double>{M[2] / bl, M[6] / bl, M[10] / bl};
    } else {
        look_at(v3(cam->eye[0], cam->eye[1], cam->eye[2]), v3(cam->target[0], cam->target[1], cam->target[2]),
                v3(cam->up[0], cam->up[1], cam->up[2]), cs->V);
        float tanHalfFovy = std::tan(cam->fovy / 2.0f);   // glm::perspective
        if (!(tanHalfFovy > 0.0f)) return false;
        float aspect = cam->aspect > 0.0f ? cam->aspect : (float)W / (float)H;
        cs->kx = tanHalfFovy * aspect;
        cs->ky = tanHalfFovy;
        cs->znear = cam->znear;
        cs->zfar = cam->zfar;
        look_at_d(cam->eye, cam->target, cam->up, cs->Vd, &cs->Sd, &cs->Ud, &cs->Bd);
        double tanHalfD = std::tan((double)cam->fovy / 2.0);
        double aspectd = cam->aspect > 0.0f ? (double)cam->aspect : (double)W / (double)H;
        cs->kxd = tanHalfD * aspectd;
        cs->kyd = tanHalfD;
    }
    cs->bb = billboard_basis(cs->V);
    return true;
}

inline uint8_t q8(float v) {  // GL float -> UNORM8 conversion (round to nearest)
    v = std::min(std::max(v, 0.0f), 1.0f);
    return (uint8_t)(int)std::floor(v * 255.0f + 0.5f);
}

// ------------------------------------------------------------------------
// Soft-particle noise march (res/conetrace_frag.glsl:103-120,136-174) and its
// noise map (src/Shaders/ConeTraceShader.cpp:100-159)
// ------------------------------------------------------------------------
constexpr int kNoiseDim = 32;

// (char)v for a float v, as the x86 conversion does it: truncate to a 32-bit
// integer (NaN and out-of-range values give INT_MIN), keep the low byte.
inline int8_t to_char(float v) {
    int32_t i = (v == v && std::fabs(v) < 2147483648.0f) ? (int32_t)v : INT32_MIN;
    return (int8_t)(uint8_t)((uint32_t)i & 0xffu);
}
// getIndex (src/Shaders/ConeTraceShader.cpp:100-113): wrap into [0, dim)
inline int noise_index(int x, int y, int z) {
    x = (x + kNoiseDim) % kNoiseDim; y = (y + kNoiseDim) % kNoiseDim; z = (z + kNoiseDim) % kNoiseDim;
    return x + y * kNoiseDim + z * kNoiseDim * kNoiseDim;
}

// initNoiseMap(32) (src/Shaders/ConeTraceShader.cpp:127-159), m[4 * i + c]
// (r, g, b, a) bytes, i = x + 32 * (y + 32 * z):
//   a = (char)genRandom(-128, 128) = (char)(u * (128 - -128) + -128)
//     (src/Util.hpp:23-33); u = lattice_value(seed, x, y, z) in [0, 1)
//     stands in for rand() / RAND_MAX;
//   gradient component = getDensity(+1) - getDensity(-1) / heightAdjust (the
//     division binds to the second term only, as written), getDensity = a / 128,
//     heightAdjust = 0.5; rgb = (char)(normalize(gradient) * 128).
// Pinned by: test_oracle_noise_map_alpha (uniform bytes over [-127, 127]),
// test_oracle_noise_map_normals (unit length; along the gradient as written,
// not along the central difference).
void noise_map_bytes(uint32_t seed, int8_t* m) {
    const int D = kNoiseDim;
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) {
                float u = lattice_value<float>(seed, x, y, z);
                m[4 * noise_index(x, y, z) + 3] = to_char(u * (128.0f - -128.0f) + -128.0f);
            }
    auto dens = [&](int i) { return (float)m[4 * i + 3] / 128.0f; };
    const float heightAdjust = 0.5f;
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) {
                V3 g = v3(dens(noise_index(x + 1, y, z)) - dens(noise_index(x - 1, y, z)) / heightAdjust,
                          dens(noise_index(x, y + 1, z)) - dens(noise_index(x, y - 1, z)) / heightAdjust,
                          dens(noise_index(x, y, z + 1)) - dens(noise_index(x, y, z - 1)) / heightAdjust);
                V3 n = normalize(g);
                int i = noise_index(x, y, z);
                m[4 * i + 0] = to_char(n.x * 128.0f);
                m[4 * i + 1] = to_char(n.y * 128.0f);
                m[4 * i + 2] = to_char(n.z * 128.0f);
            }
}

// RGBA8_SNORM texel -> float (GL 4.4 eq. 2.2): max(c / 127, -1)
template <class T> inline T snorm8(int8_t c) { return std::max((T)c / (T)127, (T)-1); }

// texture(noiseMap, s) with GL_LINEAR and GL_REPEAT on the one-level 32^3 map
// (GL 4.4 8.14.2): u = s * 32 - 1/2 as fma(s, 32, -1/2), i0 = floor(u) mod 32,
// i1 = (i0 + 1) mod 32, alpha = u - floor(u); lerps along x, then y, then z as
// in trilinear() (lerp(a, b, t) = fma(t, b - a, a)).  floor(u) mod 32 is
// taken as fl - 32 * floor(fl / 32), exact for every float fl.
// Pinned by: test_oracle_noise_sample (weight-product form, wrap-around).
template <class T> inline int wrap32(T fl) {
    T r = fl - (T)32 * std::floor(fl * (T)(1.0 / 32.0));
    int i = (r == r) ? (int)r : 0;
    return i & 31;
}
template <class T> void noise_tex(const int8_t* m, T sx, T sy, T sz, T out[4]) {
    T u[3] = {std::fma(sx, (T)32, (T)-0.5), std::fma(sy, (T)32, (T)-0.5), std::fma(sz, (T)32, (T)-0.5)};
    int i0[3], i1[3];
    T a[3];
    for (int ax = 0; ax < 3; ++ax) {
        T fl = std::floor(u[ax]);
        a[ax] = u[ax] - fl;
        i0[ax] = wrap32(fl);
        i1[ax] = (i0[ax] + 1) & 31;
    }
    auto tx = [&](int x, int y, int z, int c) { return snorm8<T>(m[4 * (x + 32 * (y + 32 * z)) + c]); };
    for (int c = 0; c < 4; ++c) {
        T c00 = lerpF(tx(i0[0], i0[1], i0[2], c), tx(i1[0], i0[1], i0[2], c), a[0]);
        T c10 = lerpF(tx(i0[0], i1[1], i0[2], c), tx(i1[0], i1[1], i0[2], c), a[0]);
        T c01 = lerpF(tx(i0[0], i0[1], i1[2], c), tx(i1[0], i0[1], i1[2], c), a[0]);
        T c11 = lerpF(tx(i0[0], i1[1], i1[2], c), tx(i1[0], i1[1], i1[2], c), a[0]);
        T c0 = lerpF(c00, c10, a[1]);
        T c1 = lerpF(c01, c11, a[1]);
        out[c] = lerpF(c0, c1, a[2]);
    }
}

template <class T> struct NV3 { T x, y, z; };
template <class T> inline T ndot(NV3<T> a, NV3<T> b) { T x = a.x * b.x, y = a.y * b.y, z = a.z * b.z; return (x + y) + z; }

// noise3D (res/conetrace_frag.glsl:103-120): octave i samples at
// (uv + octaveOffsets[i]) * freq, octaveOffsets being the vec3 uniform indexed
// by component (uv + a scalar); components past the third (GLSL leaves a
// dynamic index out of range undefined) add 0.  alpha = |sum of alphas|.
template <class T> void noise3d(const int8_t* m, const cvo_noise_march* nm, NV3<T> uv, T out[4]) {
    for (int c = 0; c < 4; ++c) out[c] = (T)0;
    T freq = (T)1, pers = (T)1;
    for (int i = 0; i < nm->octaves; ++i) {
        T off = i < 3 ? (T)nm->offsets[i] : (T)0;
        T ov[4];
        noise_tex<T>(m, (uv.x + off) * freq, (uv.y + off) * freq, (uv.z + off) * freq, ov);
        for (int c = 0; c < 4; ++c) out[c] = out[c] + pers * ov[c];
        freq = freq * (T)nm->freq_step;
        pers = pers * (T)nm->pers_step;
    }
    out[3] = std::fabs(out[3]);
}

template <class T> inline T sat(T x) { return std::fmin(std::fmax(x, (T)0), (T)1); }

// Parameters the noise march accepts: iSteps = min(len / stepSize, max - min)
// + min and the march divides by iSteps - 1 (res/conetrace_frag.glsl:154-158),
// so min = max = 1 would divide by zero; max >= 2 keeps iSteps - 1 > 0 for
// every kept fragment (its chord length is positive).
bool noise_march_valid(const cvo_noise_march* nm) {
    return nm->octaves >= 1 && nm->octaves <= 16 && nm->min_steps >= 1 && nm->max_steps >= nm->min_steps &&
           nm->max_steps <= 64 && nm->max_steps >= 2 && nm->step_size > 0.0f && nm->adjust_size != 0.0f;
}

// The noise branch of main() (res/conetrace_frag.glsl:136-174) for one
// fragment: fp = fragPos, c = center, r = radius, (tu, tv) = vertPos.xy (the
// fragment's offset in billboard units, fragTex = (vertPos.xy + 1) / 2),
// ray = normalize(V's third row).  Returns false where the fragment is
// discarded (ray misses the sphere: disc < DIST_BIAS); else col and the
// fragment alpha runningOpacity * alpha.  raySphereIntersect is :83-100.
// Pinned by: test_oracle_noise_fragment (an FP64 restatement from the GLSL
// text, with the step count, offsets and discard decisions checked).
// replay >= 0 (FP64 replay of an FP32 fragment): skip the discard test and
// take exactly `replay` steps.
template <class T>
bool noise_fragment(const int8_t* m, const cvo_noise_march* nm, NV3<T> fp, NV3<T> c, T r, T tu, T tv, NV3<T> ray,
                    T* col_out, T* alpha_out, int* steps_out, int replay = -1) {
    NV3<T> delta = {fp.x - c.x, fp.y - c.y, fp.z - c.z};
    T A = ndot(ray, ray);
    T B = (T)2 * ndot(delta, ray);
    T C = ndot(delta, delta) - r * r;
    T disc = B * B - ((T)4 * A) * C;
    if (replay < 0 && disc < (T)0.01) return false;
    T sq = std::sqrt(std::max(disc, (T)0));
    T tnear = (-B - sq) / ((T)2 * A);
    T tfar = (-B + sq) / ((T)2 * A);
    NV3<T> wn = {fp.x + ray.x * tnear, fp.y + ray.y * tnear, fp.z + ray.z * tnear};
    NV3<T> wf = {fp.x + ray.x * tfar, fp.y + ray.y * tfar, fp.z + ray.z * tfar};
    NV3<T> unit = {(wn.x - c.x) / r, (wn.y - c.y) / r, (wn.z - c.z) / r};
    T fadj = (T)1 / (T)nm->adjust_size;
    NV3<T> tn = {wn.x * fadj, wn.y * fadj, wn.z * fadj};
    NV3<T> tf = {wf.x * fadj, wf.y * fadj, wf.z * fadj};
    NV3<T> dl = {tf.x - tn.x, tf.y - tn.y, tf.z - tn.z};
    T iSteps = std::sqrt(ndot(dl, dl)) / (T)nm->step_size;
    iSteps = std::min(iSteps, (T)(nm->max_steps - nm->min_steps)) + (T)nm->min_steps;
    T den = iSteps - (T)1;
    NV3<T> dt = {dl.x / den, dl.y / den, dl.z / den};
    T opAdj = (T)nm->opacity / den;
    T lightAdj = (T)1 / den;
    T ro = (T)0, rl = (T)0;
    NV3<T> cur = tn;
    int n = 0;
    for (int i = 0; replay < 0 ? (T)i < iSteps : i < replay; ++i) {
        T nc[4];
        noise3d<T>(m, nm, cur, nc);
        T inv = (T)1 / std::sqrt(ndot(unit, unit));
        NV3<T> nx = {nc[0] + unit.x * inv, nc[1] + unit.y * inv, nc[2] + unit.z * inv};
        ro = ro + nc[3] * ((T)1 - ndot(unit, unit));
        NV3<T> up = {(T)0, (T)1, (T)0};
        rl = rl + sat(ndot(nx, up) * (T)0.5 + (T)0.5);
        cur = {cur.x + dt.x, cur.y + dt.y, cur.z + dt.z};
        unit = {unit.x + dt.x, unit.y + dt.y, unit.z + dt.z};
        ++n;
    }
    T col = (T)nm->min_color + ((T)nm->color_scale * rl) * lightAdj;
    T fx = (tu + (T)1) / (T)2 - (T)0.5, fy = (tv + (T)1) / (T)2 - (T)0.5;
    T alpha = (T)1 - std::sqrt(fx * fx + fy * fy) * (T)2;
    ro = sat(ro * opAdj);
    *col_out = col;
    *alpha_out = ro * alpha;
    if (steps_out) *steps_out = n;
    return true;
}

}  // namespace

extern "C" {

// ConeTraceShader::coneTrace (src/Shaders/ConeTraceShader.cpp:15-82) with
// the fragment stage of res/conetrace_frag.glsl:176-200 and the blend state of
// src/main.cpp:94-95:
//   per camera fragment: pos = fragPos + normalize(fragNor) * radius * sc (the
//   sphere point facing the camera, :185-186), voxelPosition =
//   calculateVoxelLerp(pos), dir = lightPos - pos (point light, :189-190) or
//   the sun direction (directional); color = vec4(indirect) (:198) in mode 0,
//   (c.rgb, a) in the front-to-back modes; clamped to [0,1].
//   composite 0: GL SRC_ALPHA / ONE_MINUS_SRC_ALPHA blend in draw order
//     (order[] from cvo_sort_boards: far to near) over the clear colour, with
//     optional UNORM8 rounding of the framebuffer after every blend;
//   composite 1: front-to-back "under" over the same fragments near to far,
//     C += (1 - A) * a * rgb, A += (1 - A) * a, stop once A >= alpha_stop.
// Only rows [row0, row1) are shaded (bounded CPU timing).  img32: W*H*4
// floats (all rows written; unshaded rows keep the clear/zero value).
// img64: FP64 replay (optional).
static int cone_trace_impl(const cvo_volume* vol, const void* levels_native, const double* levels_f64, int S,
                           const float* xyz, const float* radius, const int32_t* order, const cvo_camera* cam,
                           const cvo_light* light, const cvo_cone* cp, int W, int H, int row0, int row1,
                           int row_step, float* img32, double* img64, cvo_stats* stats, const cvo_noise_march* nm,
                           const cvo_sun_sprite* sun = nullptr) {
    if (!vol || W <= 0 || H <= 0 || !cp || cp->steps < 0 || row_step < 1) return -1;
    const bool use_nm = nm && nm->enabled;
    if (use_nm && !noise_march_valid(nm)) return -1;
    std::vector<int8_t> nmap(use_nm ? kNoiseDim * kNoiseDim * kNoiseDim * 4 : 0);
    if (use_nm) noise_map_bytes(nm->seed, nmap.data());
    Grid g = make_grid(vol);
    int N = vol->dim, L = vol->levels, C = channels(vol->format);
    Volume32 v32;
    v32.N = N; v32.L = L; v32.C = C; v32.format = vol->format; v32.data = levels_native;
    v32.off = level_offsets(N, L, C);
    Volume64 v64;
    v64.N = N; v64.L = L; v64.C = C; v64.data = levels_f64;
    v64.off = v32.off;
    const bool dbl = img64 && levels_f64;
    StepTab<float> st32 = make_steps<float>(cp, L, N);
    StepTab<double> st64 = make_steps<double>(cp, L, N);

    CamSetup cs;
    if (!camera_setup(cam, W, H, &cs)) return -1;
    // noise march view ray: normalize(vec3(V[0][2], V[1][2], V[2][2])) (res/conetrace_frag.glsl:138)
    V3 ray32 = normalize(v3(cs.V[2], cs.V[6], cs.V[10]));
    NV3<float> nray = {ray32.x, ray32.y, ray32.z};
    double rl_ = std::sqrt(cs.Vd[2] * cs.Vd[2] + cs.Vd[6] * cs.Vd[6] + cs.Vd[10] * cs.Vd[10]);
    NV3<double> nrayd = {cs.Vd[2] / rl_, cs.Vd[6] / rl_, cs.Vd[10] / rl_};

    // directional cone direction: normalize(L) / voxelDim
    V3 ld = normalize(v3(light->direction[0], light->direction[1], light->direction[2]));
    float fN = (float)N;
    Vec3T<float> dirc = {ld.x / fN, ld.y / fN, ld.z / fN};
    double lx = light->direction[0], ly = light->direction[1], lz = light->direction[2];
    double ll = std::sqrt(lx * lx + ly * ly + lz * lz);
    Vec3T<double> dircd = {lx / ll / N, ly / ll / N, lz / ll / N};

    // pixel state
    size_t npx = (size_t)W * H;
    std::vector<float> C32(npx * 4);
    std::vector<double> C64(dbl ? npx * 4 : 0);
    std::vector<uint8_t> done(npx, 0);
    std::vector<uint8_t> fb8(cp->quantize_rgba8 ? npx * 4 : 0);
    for (size_t p = 0; p < npx; ++p)
        for (int q = 0; q < 4; ++q) {
            float init = cp->composite == 0 ? cp->clear_color[q] : 0.0f;
            if (dbl) C64[4 * p + q] = init;
            if (cp->quantize_rgba8 && cp->composite == 0) {
                fb8[4 * p + q] = q8(init);
                init = (float)fb8[4 * p + q] / 255.0f;
            }
            C32[4 * p + q] = init;
        }

    // SunShader::render (src/Shaders/SunShader.cpp:7-42) after the clear and
    // before the clouds (src/main.cpp:112-116), GL blend composite only:
    // fragment at the pixel ray's hit with the view-aligned quad of half-size
    // outerRadius centred on Sun::position, dist = distance(center, fragPos)
    // (res/sun_frag.glsl:14-28), clipped by the camera's depth range.
    if (sun && sun->enabled && cp->composite == 0) {
        V3 sp = v3(light->position[0], light->position[1], light->position[2]);
        V3 sv = view_point(cs.V, sp);
        float dz = -sv.z;
        double svx = ((cs.Vd[0] * sp.x + cs.Vd[4] * sp.y) + cs.Vd[8] * sp.z) + cs.Vd[12];
        double svy = ((cs.Vd[1] * sp.x + cs.Vd[5] * sp.y) + cs.Vd[9] * sp.z) + cs.Vd[13];
        double dzd = -(((cs.Vd[2] * sp.x + cs.Vd[6] * sp.y) + cs.Vd[10] * sp.z) + cs.Vd[14]);
        const float r0 = sun->inner_radius, r1 = sun->outer_radius;
        if (dz >= cs.znear && dz <= cs.zfar) {
            for (int py = std::max(row0, 0); py < std::min(row1, H); ++py) {
                float ry = pixel_ndc(py, H) * cs.ky;
                double ryd = ((2.0 * ((double)py + 0.5)) / (double)H - 1.0) * cs.kyd;
                for (int px = 0; px < W; ++px) {
                    float rx = pixel_ndc(px, W) * cs.kx;
                    float du = rx * dz - sv.x, dv = ry * dz - sv.y;
                    if (!(std::fabs(du) <= r1 && std::fabs(dv) <= r1)) continue;
                    float dist = std::sqrt(du * du + dv * dv);
                    float q[4];
                    if (dist < r0) {
                        for (int c = 0; c < 3; ++c) q[c] = sun->inner_color[c];
                        q[3] = 1.0f;
                    } else {
                        float sc = (dist - r0) / (r1 - r0);
                        if (sc > 0.99f) continue;
                        for (int c = 0; c < 3; ++c) q[c] = sun->outer_color[c] * sc + sun->inner_color[c] * (1.0f - sc);
                        q[3] = 1.0f - sc;
                    }
                    size_t pix = (size_t)py * W + px;
                    float a = sat(q[3]);
                    for (int c = 0; c < 4; ++c) {
                        float dst = cp->quantize_rgba8 ? (float)fb8[4 * pix + c] / 255.0f : C32[4 * pix + c];
                        float res = sat(q[c]) * a + dst * (1.0f - a);
                        if (cp->quantize_rgba8) {
                            fb8[4 * pix + c] = q8(res);
                            C32[4 * pix + c] = (float)fb8[4 * pix + c] / 255.0f;
                        } else {
                            C32[4 * pix + c] = res;
                        }
                    }
                    if (dbl) {
                        // FP64 replay of the same (FP32-decided) fragment
                        double rxd = ((2.0 * ((double)px + 0.5)) / (double)W - 1.0) * cs.kxd;
                        double dud = rxd * dzd - svx, dvd = ryd * dzd - svy;
                        double dd = std::sqrt(dud * dud + dvd * dvd);
                        double qd[4];
                        if (dist < r0) {
                            for (int c = 0; c < 3; ++c) qd[c] = sun->inner_color[c];
                            qd[3] = 1.0;
                        } else {
                            double sc = (dd - r0) / ((double)r1 - r0);
                            for (int c = 0; c < 3; ++c)
                                qd[c] = (double)sun->outer_color[c] * sc + (double)sun->inner_color[c] * (1.0 - sc);
                            qd[3] = 1.0 - sc;
                        }
                        double ad = sat(qd[3]);
                        for (int c = 0; c < 4; ++c) C64[4 * pix + c] = sat(qd[c]) * ad + C64[4 * pix + c] * (1.0 - ad);
                    }
                }
            }
        }
    }

    // per-sphere setup in draw order
    struct Sph { V3 c, cv; float s; PixBox box; bool live; };
    std::vector<Sph> sph(S);
    for (int r = 0; r < S; ++r) {
        int j = order ? order[r] : r;
        Sph& o = sph[r];
        o.c = v3(vol->position[0] + xyz[3 * j], vol->position[1] + xyz[3 * j + 1], vol->position[2] + xyz[3 * j + 2]);
        o.s = radius[j];
        o.cv = view_point(cs.V, o.c);
        float dz = -o.cv.z;
        o.live = (o.s > 0.0f) && (dz >= cs.znear && dz <= cs.zfar);
        if (o.live) {
            o.box = pixel_box(o.cv.x, o.cv.y, o.s, cs.kx * dz, cs.ky * dz, 0.0f, 0.0f, W, H);
            o.live = !o.box.empty;
        }
    }

    int64_t nfrag = 0, nsamp = 0, ntwo = 0;
    int r0 = std::max(row0, 0), r1 = std::min(row1, H);
    // count two-level samples per step
    std::vector<int> two(cp->steps, 0);
    for (int i = 0; i < cp->steps; ++i) two[i] = st32.frac[i] != 0.0f ? 1 : 0;

#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nfrag, nsamp, ntwo)
    for (int py = r0; py < r1; ++py) {
        if ((py - r0) % row_step != 0) continue;
        float yn = pixel_ndc(py, H);
        float ry = yn * cs.ky;
        double ynd = (2.0 * ((double)py + 0.5)) / (double)H - 1.0;
        double ryd = ynd * cs.kyd;
        for (int t = 0; t < S; ++t) {
            int r = cp->composite == 1 ? (S - 1 - t) : t;  // front-to-back walks near to far
            const Sph& o = sph[r];
            if (!o.live || py < o.box.y0 || py > o.box.y1) continue;
            float dz = -o.cv.z;
            float hy = ry * dz;
            float dv = hy - o.cv.y;
            int j = order ? order[r] : r;
            for (int px = o.box.x0; px <= o.box.x1; ++px) {
                size_t pix = (size_t)py * W + px;
                if (done[pix]) continue;
                float xn = pixel_ndc(px, W);
                float rx = xn * cs.kx;
                float hx = rx * dz;
                float du = hx - o.cv.x;
                float sc;
                if (!sphere_contrib(du, dv, o.s, &sc)) continue;
                // fragment position, FP32
                V3 fp = v3((o.c.x + du * cs.bb.S.x) + dv * cs.bb.U.x, (o.c.y + du * cs.bb.S.y) + dv * cs.bb.U.y,
                           (o.c.z + du * cs.bb.S.z) + dv * cs.bb.U.z);
                // noise branch first (it may discard the fragment)
                float nm_col = 1.0f, nm_alpha = 0.0f;
                int nm_steps = 0;
                if (use_nm && !noise_fragment<float>(nmap.data(), nm, NV3<float>{fp.x, fp.y, fp.z},
                                                     NV3<float>{o.c.x, o.c.y, o.c.z}, o.s, du / o.s, dv / o.s, nray,
                                                     &nm_col, &nm_alpha, &nm_steps))
                    continue;
                ++nfrag;
                // fragPos + billboardNormal * radius * sphereContrib: (n * r) * sc
                V3 pos = v3(fp.x + (cs.bb.B.x * o.s) * sc, fp.y + (cs.bb.B.y * o.s) * sc, fp.z + (cs.bb.B.z * o.s) * sc);
                V3 vp = voxel_lerp(g, pos);
                Vec3T<float> p0 = {vp.x / fN, vp.y / fN, vp.z / fN};
                Vec3T<float> d;
                if (light->directional) {
                    d = dirc;
                } else {
                    V3 dd = normalize(sub(v3(light->position[0], light->position[1], light->position[2]), pos));
                    d.x = dd.x / fN; d.y = dd.y / fN; d.z = dd.z / fN;
                }
                ConeResult<float> cr = trace_cone<float>(v32, st32, cp, p0, d, -1);
                nsamp += cr.nsamp;
                for (int i = 0; i < cr.nsamp; ++i) ntwo += two[i];
                float src[4];
                for (int q = 0; q < 4; ++q) src[q] = std::min(std::max(cr.rgba[q], 0.0f), 1.0f);
                if (use_nm) {
                    // color = vec4(vec3(col), runningOpacity * alpha); color.rgb *= indirect (:172, :194-196)
                    for (int q = 0; q < 3; ++q) src[q] = sat(nm_col * cr.rgba[q]);
                    src[3] = sat(nm_alpha);
                }

                // FP64 replay of the same fragment
                double srcd[4] = {0, 0, 0, 0};
                if (dbl) {
                    double xnd = (2.0 * ((double)px + 0.5)) / (double)W - 1.0;
                    double dzd = -(((cs.Vd[2] * o.c.x + cs.Vd[6] * o.c.y) + cs.Vd[10] * o.c.z) + cs.Vd[14]);
                    double cvx = ((cs.Vd[0] * o.c.x + cs.Vd[4] * o.c.y) + cs.Vd[8] * o.c.z) + cs.Vd[12];
                    double cvy = ((cs.Vd[1] * o.c.x + cs.Vd[5] * o.c.y) + cs.Vd[9] * o.c.z) + cs.Vd[13];
                    double dud = xnd * cs.kxd * dzd - cvx, dvd = ryd * dzd - cvy;
                    double qd = std::sqrt(dud * dud + dvd * dvd) / (double)o.s;
                    double scd = std::sqrt(std::max(0.0, 1.0 - qd * qd));
                    double px_ = o.c.x + dud * cs.Sd.x + dvd * cs.Ud.x + cs.Bd.x * o.s * scd;
                    double py_ = o.c.y + dud * cs.Sd.y + dvd * cs.Ud.y + cs.Bd.y * o.s * scd;
                    double pz_ = o.c.z + dud * cs.Sd.z + dvd * cs.Ud.z + cs.Bd.z * o.s * scd;
                    double lox = (double)vol->position[0] + (double)vol->bounds_lo[0];
                    double loy = (double)vol->position[1] + (double)vol->bounds_lo[1];
                    double loz = (double)vol->position[2] + (double)vol->bounds_lo[2];
                    double rxr = ((double)vol->position[0] + vol->bounds_hi[0]) - lox;
                    double ryr = ((double)vol->position[1] + vol->bounds_hi[1]) - loy;
                    double rzr = ((double)vol->position[2] + vol->bounds_hi[2]) - loz;
                    Vec3T<double> p0d = {(px_ - lox) / rxr, (py_ - loy) / ryr, (pz_ - loz) / rzr};
                    Vec3T<double> dd;
                    if (light->directional) dd = dircd;
                    else {
                        double ax = light->position[0] - px_, ay = light->position[1] - py_, az = light->position[2] - pz_;
                        double al = std::sqrt(ax * ax + ay * ay + az * az);
                        dd.x = ax / al / N; dd.y = ay / al / N; dd.z = az / al / N;
                    }
                    ConeResult<double> crd = trace_cone<double>(v64, st64, cp, p0d, dd, cr.nsamp);
                    for (int q = 0; q < 4; ++q) srcd[q] = std::min(std::max(crd.rgba[q], 0.0), 1.0);
                    if (use_nm) {
                        double fpx = o.c.x + dud * cs.Sd.x + dvd * cs.Ud.x;
                        double fpy = o.c.y + dud * cs.Sd.y + dvd * cs.Ud.y;
                        double fpz = o.c.z + dud * cs.Sd.z + dvd * cs.Ud.z;
                        double cold = 1.0, alphad = 0.0;
                        noise_fragment<double>(nmap.data(), nm, NV3<double>{fpx, fpy, fpz},
                                               NV3<double>{o.c.x, o.c.y, o.c.z}, (double)o.s, dud / o.s, dvd / o.s,
                                               nrayd, &cold, &alphad, nullptr, nm_steps);
                        for (int q = 0; q < 3; ++q) srcd[q] = sat(cold * crd.rgba[q]);
                        srcd[3] = sat(alphad);
                    }
                }
                (void)j;

                if (cp->composite == 0) {
                    // dst = src * src.a + dst * (1 - src.a) on all four channels
                    float a = src[3];
                    for (int q = 0; q < 4; ++q) {
                        float dst = cp->quantize_rgba8 ? (float)fb8[4 * pix + q] / 255.0f : C32[4 * pix + q];
                        float res = src[q] * a + dst * (1.0f - a);
                        if (cp->quantize_rgba8) {
                            fb8[4 * pix + q] = q8(res);
                            C32[4 * pix + q] = (float)fb8[4 * pix + q] / 255.0f;
                        } else {
                            C32[4 * pix + q] = res;
                        }
                    }
                    if (dbl) {
                        double ad = srcd[3];
                        for (int q = 0; q < 4; ++q) C64[4 * pix + q] = srcd[q] * ad + C64[4 * pix + q] * (1.0 - ad);
                    }
                } else {
                    float A = C32[4 * pix + 3];
                    float w = (1.0f - A) * src[3];
                    for (int q = 0; q < 3; ++q) C32[4 * pix + q] = std::fma(w, src[q], C32[4 * pix + q]);
                    C32[4 * pix + 3] = A + w;
                    if (dbl) {
                        double Ad = C64[4 * pix + 3];
                        double wd = (1.0 - Ad) * srcd[3];
                        for (int q = 0; q < 3; ++q) C64[4 * pix + q] = std::fma(wd, srcd[q], C64[4 * pix + q]);
                        C64[4 * pix + 3] = Ad + wd;
                    }
                    if (C32[4 * pix + 3] >= cp->alpha_stop) done[pix] = 1;
                }
            }
        }
    }
    std::memcpy(img32, C32.data(), npx * 4 * sizeof(float));
    if (dbl) std::memcpy(img64, C64.data(), npx * 4 * sizeof(double));
    if (stats) { stats->fragments = nfrag; stats->samples = nsamp; stats->two_level = ntwo; }
    return 0;
}

int cvo_cone_trace(const cvo_volume* vol, const void* levels_native, const double* levels_f64, int S,
                   const float* xyz, const float* radius, const int32_t* order, const cvo_camera* cam,
                   const cvo_light* light, const cvo_cone* cp, int W, int H, int row0, int row1, int row_step,
                   float* img32, double* img64, cvo_stats* stats) {
    return cone_trace_impl(vol, levels_native, levels_f64, S, xyz, radius, order, cam, light, cp, W, H, row0, row1,
                           row_step, img32, img64, stats, nullptr);
}

// cvo_cone_trace with the soft-particle noise branch of res/conetrace_frag.glsl
// (:136-174, :194-196) in front of the cone: a fragment the noise branch
// discards is dropped; otherwise its colour is (col * indirect rgb,
// runningOpacity * alpha), clamped to [0, 1].
// cvo_cone_trace_nm with the sun sprite drawn first (sun may be NULL)
int cvo_cone_trace_ex(const cvo_volume* vol, const void* levels_native, const double* levels_f64, int S,
                      const float* xyz, const float* radius, const int32_t* order, const cvo_camera* cam,
                      const cvo_light* light, const cvo_cone* cp, int W, int H, int row0, int row1, int row_step,
                      float* img32, double* img64, cvo_stats* stats, const cvo_noise_march* nm,
                      const cvo_sun_sprite* sun) {
    return cone_trace_impl(vol, levels_native, levels_f64, S, xyz, radius, order, cam, light, cp, W, H, row0, row1,
                           row_step, img32, img64, stats, nm, sun);
}

int cvo_cone_trace_nm(const cvo_volume* vol, const void* levels_native, const double* levels_f64, int S,
                      const float* xyz, const float* radius, const int32_t* order, const cvo_camera* cam,
                      const cvo_light* light, const cvo_cone* cp, int W, int H, int row0, int row1, int row_step,
                      float* img32, double* img64, cvo_stats* stats, const cvo_noise_march* nm) {
    return cone_trace_impl(vol, levels_native, levels_f64, S, xyz, radius, order, cam, light, cp, W, H, row0, row1,
                           row_step, img32, img64, stats, nm);
}

// the noise map bytes (32^3 x (r, g, b, a), x fastest)
void cvo_noise_map(uint32_t seed, int8_t* out) { noise_map_bytes(seed, out); }

// texture(noiseMap, p) at n points (FP32 and FP64), for tests
void cvo_noise_sample(uint32_t seed, const float* pts, int64_t n, float* out32, double* out64) {
    std::vector<int8_t> m(kNoiseDim * kNoiseDim * kNoiseDim * 4);
    noise_map_bytes(seed, m.data());
    for (int64_t i = 0; i < n; ++i) {
        noise_tex<float>(m.data(), pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], out32 + 4 * i);
        if (out64) noise_tex<double>(m.data(), pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], out64 + 4 * i);
    }
}

// one fragment of the noise branch (FP32), for tests: in[0..2] fragPos,
// in[3..5] center, in[6] radius, in[7..8] vertPos.xy, in[9..11] view ray;
// out = (col, alpha, steps, kept)
int cvo_noise_march_valid(const cvo_noise_march* nm) { return noise_march_valid(nm) ? 1 : 0; }

// The camera as the cone trace sets it up, for tests: out[0..15] V,
// out[16..17] the projection scales (kx, ky), out[18..19] znear, zfar,
// out[20..28] the billboard basis S, U, B, out[29..31] the noise march's
// view ray normalize(V[0][2], V[1][2], V[2][2]) (res/conetrace_frag.glsl:138).
// Returns -1 on an unusable camera.
int cvo_camera_setup(const cvo_camera* cam, int W, int H, float* out) {
    CamSetup cs;
    if (!camera_setup(cam, W, H, &cs)) return -1;
    for (int i = 0; i < 16; ++i) out[i] = cs.V[i];
    out[16] = cs.kx; out[17] = cs.ky; out[18] = cs.znear; out[19] = cs.zfar;
    const V3 b[3] = {cs.bb.S, cs.bb.U, cs.bb.B};
    for (int k = 0; k < 3; ++k) { out[20 + 3 * k] = b[k].x; out[21 + 3 * k] = b[k].y; out[22 + 3 * k] = b[k].z; }
    V3 ray = normalize(v3(cs.V[2], cs.V[6], cs.V[10]));
    out[29] = ray.x; out[30] = ray.y; out[31] = ray.z;
    return 0;
}

void cvo_noise_fragment(const cvo_noise_march* nm, const float* in, float* out) {
    std::vector<int8_t> m(kNoiseDim * kNoiseDim * kNoiseDim * 4);
    noise_map_bytes(nm->seed, m.data());
    float col = 0.0f, alpha = 0.0f;
    int steps = 0;
    bool kept = noise_fragment<float>(m.data(), nm, NV3<float>{in[0], in[1], in[2]}, NV3<float>{in[3], in[4], in[5]},
                                      in[6], in[7], in[8], NV3<float>{in[9], in[10], in[11]}, &col, &alpha, &steps);
    out[0] = col; out[1] = alpha; out[2] = (float)steps; out[3] = kept ? 1.0f : 0.0f;
}

// the FP32 cone schedule, for tests: h_i, r_i, level_i, frac_i
int cvo_step_table(const cvo_cone* cp, int levels, int dim, float* h, float* r, int32_t* lvl, float* frac) {
    StepTab<float> t = make_steps<float>(cp, levels, dim);
    for (int i = 0; i < cp->steps; ++i) { h[i] = t.h[i]; r[i] = t.r[i]; lvl[i] = t.lvl[i]; frac[i] = t.frac[i]; }
    return 0;
}

// direct access to the sampler and the cone for known-answer tests
int cvo_sample(const cvo_volume* vol, const void* levels_native, const double* levels_f64, float lam, const float* pts,
               int64_t n, float* out32, double* out64) {
    int N = vol->dim, L = vol->levels, C = channels(vol->format);
    Volume32 v32;
    v32.N = N; v32.L = L; v32.C = C; v32.format = vol->format; v32.data = levels_native;
    v32.off = level_offsets(N, L, C);
    Volume64 v64;
    v64.N = N; v64.L = L; v64.C = C; v64.data = levels_f64; v64.off = v32.off;
    StepTab<float> st;
    StepTab<double> sd;
    int lv; float f;
    if (lam <= 0.0f) { lv = 0; f = 0.0f; }
    else if (lam >= (float)(L - 1)) { lv = L - 1; f = 0.0f; }
    else { float fl = std::floor(lam); lv = (int)fl; f = lam - fl; }
    st.lvl.push_back(lv); st.frac.push_back(f);
    sd.lvl.push_back(lv); sd.frac.push_back((double)f);
    static const int chs[4] = {0, 1, 2, 3};
    for (int64_t i = 0; i < n; ++i) {
        sample_lod<float>(v32, st, 0, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], C, chs, out32 + C * i);
        if (out64 && levels_f64)
            sample_lod<double>(v64, sd, 0, (double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2], C, chs,
                               out64 + C * i);
    }
    return 0;
}

int cvo_trace_one(const cvo_volume* vol, const void* levels_native, const double* levels_f64, const cvo_cone* cp,
                  const float p0[3], const float d[3], float* out32, double* out64, int32_t* nsamp) {
    int N = vol->dim, L = vol->levels, C = channels(vol->format);
    Volume32 v32;
    v32.N = N; v32.L = L; v32.C = C; v32.format = vol->format; v32.data = levels_native;
    v32.off = level_offsets(N, L, C);
    Volume64 v64;
    v64.N = N; v64.L = L; v64.C = C; v64.data = levels_f64; v64.off = v32.off;
    StepTab<float> st32 = make_steps<float>(cp, L, N);
    StepTab<double> st64 = make_steps<double>(cp, L, N);
    Vec3T<float> p = {p0[0], p0[1], p0[2]}, dd = {d[0], d[1], d[2]};
    ConeResult<float> r = trace_cone<float>(v32, st32, cp, p, dd, -1);
    for (int q = 0; q < 4; ++q) out32[q] = r.rgba[q];
    *nsamp = r.nsamp;
    if (out64 && levels_f64) {
        Vec3T<double> pd = {p0[0], p0[1], p0[2]}, d2 = {d[0], d[1], d[2]};
        ConeResult<double> rd = trace_cone<double>(v64, st64, cp, pd, d2, r.nsamp);
        for (int q = 0; q < 4; ++q) out64[q] = rd.rgba[q];
    }
    return 0;
}

int cvo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// threads of the following parallel regions (n >= 1)
void cvo_set_num_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"

// ========================================================================
// learned far field: segment features / targets and the MLP
// ========================================================================
// No reference code exists for this stage (it is defined by this project);
// the functions below restate the HIP kernels' definitions so that the GPU
// path can be checked, and are pinned by closed forms (features on constant
// volumes), the plain cone march (targets) and finite differences (MLP
// gradients).
namespace {

// tanHalf and the feature sampling distances: the cone radius equals
// 2^(lod-1) voxels (lambda = lod) at height 2^(lod-1) / tan(angle / 2);
// LODs 3..6, clamped to the last level.
struct FarField {
    int k;
    float feat_h[4];
    int feat_lvl[4];
};

FarField make_farfield(const cvo_cone* cp, int L, int k) {
    FarField f;
    f.k = k;
    float tanHalf = std::tan(cp->cone_angle / 2.0f);
    for (int j = 0; j < 4; ++j) {
        int lod = 3 + j;
        f.feat_h[j] = (float)(1 << (lod - 1)) / tanHalf;
        f.feat_lvl[j] = std::min(lod, L - 1);
    }
    return f;
}

// 17 features after k samples: last sample position (2x - 1), cone direction
// (d * N), sun direction, then (density, light) at LODs 3..6 at distance
// max(h_k, feat_h[j]) along the cone ((0, 0) outside [0,1]^3).
void farfield_features(const Volume32& v, const FarField& ff, const StepTab<float>& st, const float sun[3],
                       Vec3T<float> p, Vec3T<float> d, float fN, float* f) {
    float hk = st.h[ff.k - 1];
    float xx = std::fma(hk, d.x, p.x), xy = std::fma(hk, d.y, p.y), xz = std::fma(hk, d.z, p.z);
    f[0] = 2.0f * xx - 1.0f;
    f[1] = 2.0f * xy - 1.0f;
    f[2] = 2.0f * xz - 1.0f;
    f[3] = d.x * fN;
    f[4] = d.y * fN;
    f[5] = d.z * fN;
    f[6] = sun[0];
    f[7] = sun[1];
    f[8] = sun[2];
    static const int chs[2] = {3, 0};
    for (int j = 0; j < 4; ++j) {
        float hj = std::max(hk, ff.feat_h[j]);
        float sx = std::fma(hj, d.x, p.x), sy = std::fma(hj, d.y, p.y), sz = std::fma(hj, d.z, p.z);
        float o[2] = {0.0f, 0.0f};
        if (sx >= 0.0f && sx <= 1.0f && sy >= 0.0f && sy <= 1.0f && sz >= 0.0f && sz <= 1.0f)
            trilinear<float>(v, ff.feat_lvl[j], sx, sy, sz, 2, chs, o);
        f[9 + 2 * j] = o[0];
        f[10 + 2 * j] = o[1];
    }
}

// ---- MLP 17 -> 64 -> 64 -> 64 -> 64 -> 5 (flat FP32 parameters:
// W0 [64][17], b0, W1 [64][64], b1, W2, b2, W3, b3, W4 [5][64], b4)
constexpr int kIn = 17, kHid = 64, kOut = 5;
inline int lin(int l) { return l == 0 ? kIn : kHid; }
inline int lout(int l) { return l == 4 ? kOut : kHid; }
inline int woff(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += lin(i) * lout(i) + lout(i);
    return o;
}
inline int boff(int l) { return woff(l) + lin(l) * lout(l); }
constexpr int kParams = 13957;

// bf16 round to nearest even (the GPU's conversion; NaN kept quiet)
inline float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) u = ((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u)) << 16;
    else u = ((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
inline double rnd(double x, int emulate) { return emulate ? (double)bf16r((float)x) : x; }

// GELU, tanh form: x * sigmoid(2 sqrt(2/pi) (x + 0.044715 x^3))
constexpr double kC1 = 1.5957691216057308, kC2 = 1.5957691216057308 * 0.044715;
inline double gelu_d(double x) { return x / (1.0 + std::exp(-x * (kC1 + kC2 * x * x))); }
inline double gelu_grad_d(double x) {
    double sg = 1.0 / (1.0 + std::exp(-x * (kC1 + kC2 * x * x)));
    return sg + x * sg * (1.0 - sg) * (kC1 + 3.0 * kC2 * x * x);
}

struct MlpTrace {
    double H[5][kHid];  // layer inputs (rounded as the GPU rounds them)
    double Z[4][kHid];  // hidden pre-activations
    double y[kOut];
};

// emulate: 1 = round the matrix-core operands (inputs, weights, activations)
// to bf16 as the GPU does; 0 = pure FP64 network.
void mlp_fwd(const float* P, const float* x, int emulate, MlpTrace& t) {
    for (int k = 0; k < kIn; ++k) t.H[0][k] = rnd((double)x[k], emulate);
    for (int l = 0; l < 5; ++l) {
        const float* W = P + woff(l);
        const float* b = P + boff(l);
        for (int o = 0; o < lout(l); ++o) {
            double z = (double)b[o];
            for (int k = 0; k < lin(l); ++k) z += rnd((double)W[o * lin(l) + k], emulate) * t.H[l][k];
            if (l < 4) {
                t.Z[l][o] = z;
                t.H[l + 1][o] = rnd(gelu_d(z), emulate);
            } else {
                t.y[o] = z;
            }
        }
    }
}

}  // namespace

extern "C" {

// Segment samples of the learned far field.  For each start point p0 (texture
// space) the front-to-back cone (modes 1/2, directional light) is marched as
// in cvo_cone_trace; if it is still running after k samples (not exited, not
// saturated, k < steps) the 17 features are recorded with the state (c_k, a_k),
// the march continues to its end (c, a), and the targets are the far field in
// units of the remaining transmittance:
//     t_rgb = (c - c_k) / (1 - a_k),  t_a = (a - a_k) / (1 - a_k),  t_4 = 1 - a_k.
// valid[i] = 1 for recorded rows (0 otherwise, features/targets zero).
int cvo_segments(const cvo_volume* vol, const void* levels_native, const cvo_cone* cp, const cvo_light* light, int k,
                 const float* p0, int64_t n, float* feat, float* tgt, int32_t* valid) {
    if (!vol || !cp || cp->cone_mode == 0 || k < 1 || !light || !light->directional) return -1;
    int N = vol->dim, L = vol->levels, C = channels(vol->format);
    Volume32 v32;
    v32.N = N; v32.L = L; v32.C = C; v32.format = vol->format; v32.data = levels_native;
    v32.off = level_offsets(N, L, C);
    StepTab<float> st = make_steps<float>(cp, L, N);
    FarField ff = make_farfield(cp, L, k);
    V3 ld = normalize(v3(light->direction[0], light->direction[1], light->direction[2]));
    float fN = (float)N;
    Vec3T<float> d = {ld.x / fN, ld.y / fN, ld.z / fN};
    float sun[3] = {ld.x, ld.y, ld.z};
    static const int chRGBA[4] = {0, 1, 2, 3};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Vec3T<float> p = {p0[3 * i], p0[3 * i + 1], p0[3 * i + 2]};
        float c[3] = {0, 0, 0}, a = 0, cc[3] = {0, 0, 0}, ca = 0;
        float ck[3] = {0, 0, 0}, ak = 0;
        bool rec = false;
        float f[17];
        int nsmp = 0;
        for (int s = 1; s <= cp->steps; ++s) {
            float h = st.h[s - 1];
            float sx = std::fma(h, d.x, p.x), sy = std::fma(h, d.y, p.y), sz = std::fma(h, d.z, p.z);
            if (!(sx >= 0.0f && sx <= 1.0f && sy >= 0.0f && sy <= 1.0f && sz >= 0.0f && sz <= 1.0f)) break;
            float sm[4];
            sample_lod<float>(v32, st, s - 1, sx, sy, sz, 4, chRGBA, sm);
            float al = cp->cone_mode == 1 ? sm[3] : 1.0f - cv_exp2_neg(-(st.kk[s - 1] * sm[3]));
            float w = (1.0f - a) * al;
            if (cp->compensated) {
                for (int q = 0; q < 3; ++q) {
                    float y = std::fma(w, sm[q], -cc[q]);
                    float t = c[q] + y;
                    cc[q] = (t - c[q]) - y;
                    c[q] = t;
                }
                float y = w - ca;
                float t = a + y;
                ca = (t - a) - y;
                a = t;
            } else {
                for (int q = 0; q < 3; ++q) c[q] = std::fma(w, sm[q], c[q]);
                a = a + w;
            }
            ++nsmp;
            if (a >= cp->alpha_stop) break;
            if (nsmp == k && k < cp->steps) {
                rec = true;
                ck[0] = c[0]; ck[1] = c[1]; ck[2] = c[2]; ak = a;
                farfield_features(v32, ff, st, sun, p, d, fN, f);
            }
        }
        valid[i] = rec ? 1 : 0;
        for (int q = 0; q < 17; ++q) feat[17 * i + q] = rec ? f[q] : 0.0f;
        if (rec) {
            float om = 1.0f - ak;
            tgt[5 * i + 0] = (c[0] - ck[0]) / om;
            tgt[5 * i + 1] = (c[1] - ck[1]) / om;
            tgt[5 * i + 2] = (c[2] - ck[2]) / om;
            tgt[5 * i + 3] = (a - ak) / om;
            tgt[5 * i + 4] = om;
        } else {
            for (int q = 0; q < 5; ++q) tgt[5 * i + q] = 0.0f;
        }
    }
    return 0;
}

int cvo_mlp_param_count(void) { return kParams; }

// Batched MLP inference: X [n][17] -> Y [n][5].
int cvo_mlp_forward(const float* params, const float* X, int64_t n, int emulate, float* Y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        MlpTrace t;
        mlp_fwd(params, X + 17 * i, emulate, t);
        for (int q = 0; q < kOut; ++q) Y[kOut * i + q] = (float)t.y[q];
    }
    return 0;
}

// Loss and its gradient over a batch (mean over the n rows):
//     w^2 sum_c (y_c - t_c)^2 + (y_4 - e)^2,  w = t_4,
//     e = max_c |clamp(y_c, 0, 1) - t_c| (a constant; e_fixed[i] if given).
// With emulate = 1 the backward rounds its matrix-core operands (output
// gradients, hidden gradients) to bf16 where the GPU does; bias gradients use
// the unrounded values.
int cvo_mlp_grad(const float* params, const float* X, const float* T, int64_t n, int emulate, const float* e_fixed,
                 double* grad, double* loss) {
    std::vector<double> g(kParams, 0.0);
    double lsum = 0.0;
    const double inv_n = 1.0 / (double)n;
    for (int64_t i = 0; i < n; ++i) {
        MlpTrace t;
        mlp_fwd(params, X + 17 * i, emulate, t);
        const float* tt = T + 5 * i;
        double w = tt[4];
        double e = 0.0;
        for (int q = 0; q < 4; ++q) e = std::max(e, std::fabs(std::min(std::max(t.y[q], 0.0), 1.0) - (double)tt[q]));
        if (e_fixed) e = e_fixed[i];
        double G[kHid], Gr[kHid];
        for (int q = 0; q < 4; ++q) {
            double dq = t.y[q] - (double)tt[q];
            G[q] = 2.0 * w * w * dq * inv_n;
            lsum += w * w * dq * dq;
        }
        G[4] = 2.0 * (t.y[4] - e) * inv_n;
        lsum += (t.y[4] - e) * (t.y[4] - e);
        for (int l = 4; l >= 0; --l) {
            const float* W = params + woff(l);
            for (int o = 0; o < lout(l); ++o) Gr[o] = rnd(G[o], emulate);
            for (int o = 0; o < lout(l); ++o) {
                g[boff(l) + o] += G[o];
                for (int k = 0; k < lin(l); ++k) g[woff(l) + o * lin(l) + k] += Gr[o] * t.H[l][k];
            }
            if (l > 0) {
                double Gn[kHid];
                for (int k = 0; k < lin(l); ++k) {
                    double dh = 0.0;
                    for (int o = 0; o < lout(l); ++o) dh += rnd((double)W[o * lin(l) + k], emulate) * Gr[o];
                    Gn[k] = dh * gelu_grad_d(t.Z[l - 1][k]);
                }
                for (int k = 0; k < lin(l); ++k) G[k] = Gn[k];
            }
        }
    }
    for (int p = 0; p < kParams; ++p) grad[p] = g[p];
    *loss = lsum * inv_n;
    return 0;
}

}  // extern "C"
