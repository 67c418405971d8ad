/*
 * cloudvox.h - C ABI of the MI355X (gfx950) billboard-cloud hot path:
 * billboard-sphere voxelization (two passes) into a 3D volume, the volume's
 * 3D mip chain, and the per-pixel voxel cone trace toward the sun with
 * compositing, rendered headless into a caller-owned device image.
 *
 * It replaces the GL render passes of Cloud-Renderer:
 *   VoxelizeShader::voxelize(CloudVolume*)   src/Shaders/VoxelizeShader.hpp:18,
 *                                            src/Shaders/VoxelizeShader.cpp:18-134
 *   ConeTraceShader::coneTrace(CloudVolume*) src/Shaders/ConeTraceShader.hpp:12,
 *                                            src/Shaders/ConeTraceShader.cpp:15-98
 * and the GL objects of CloudVolume that they touch
 *   CloudVolume(dim, bounds, position, mips) src/CloudVolume.cpp:7-55
 *   CloudVolume::uploadBillboards             src/CloudVolume.cpp:139-164
 *   CloudVolume::sortBoards                   src/CloudVolume.cpp:65-82
 *   Sun::update                               src/Sun.hpp:26-43
 *
 * Conventions
 *   - Every entry point returns an int status (CR_OK = 0, negative on error)
 *     and never aborts; cr_last_error() describes the last failure on the
 *     context.  No C++ exception crosses this boundary.
 *   - Host arrays passed in are copied before the call returns.
 *   - All device work is ordered on the caller's hipStream_t (passed as
 *     void*, NULL = the null stream) and is complete after that stream is
 *     synchronised (cr_voxelize and cr_cone_trace run part of it on streams
 *     of the volume that the caller's stream waits for).  Calls on one
 *     volume must not overlap.
 *   - The library owns the volume, its mips, bitmaps and scratch buffers; the
 *     caller owns the output image.
 *   - Images are stored bottom row first (GL window coordinates).
 */
#ifndef CLOUDVOX_H_
#define CLOUDVOX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CR_ABI_VERSION 3

/* status codes */
#define CR_OK 0
#define CR_EINVAL (-1)       /* invalid argument */
#define CR_ENOMEM (-2)       /* device or host allocation failed */
#define CR_EDEVICE (-3)      /* HIP runtime error / no device */
#define CR_ESTATE (-4)       /* call out of order (e.g. cone trace before voxelize) */
#define CR_EUNSUPPORTED (-5) /* valid but unsupported combination */

/* volume storage format (the reference stores GL_R8, src/CloudVolume.cpp:18) */
#define CR_FORMAT_R8 0
#define CR_FORMAT_RGBA32F 1
#define CR_FORMAT_RGBA16F 2

/* volume content mode */
#define CR_MODE_REF 0  /* shipped shaders: R = 1 on the lit shell only */
#define CR_MODE_RGBA 1 /* paper semantics: chord fill (0,0,0,1) + lit shell (1,1,1,1) */

/* cone accumulation */
#define CR_CONE_WEIGHTED 0    /* res/conetrace_frag.glsl:64-79: sum s.r * i / (steps * downScaling) */
#define CR_CONE_FTB 1         /* front-to-back, alpha_i = s.a */
#define CR_CONE_FTB_DENSITY 2 /* front-to-back, alpha_i = 1 - 2^(-sigma*log2(e)*coneRadius_i/N * s.a) */

/* pixel compositing */
#define CR_COMPOSITE_GL_BLEND 0 /* src/main.cpp:94-95 blend, back to front, all fragments */
#define CR_COMPOSITE_FTB 1      /* front to back, early-out at alpha_stop */

/* output pixel format */
#define CR_PIXEL_RGBA32F 0
#define CR_PIXEL_RGBA8 1

typedef struct cr_context_s *cr_context;
typedef struct cr_volume_s *cr_volume;
typedef struct cr_mlp_s *cr_mlp;
typedef struct cr_batch_s *cr_batch;

/* Sun (src/Sun.hpp:12-19, src/main.cpp:37).  position drives the light view
 * of the voxelization exactly as Sun::update does; the cone direction is
 * (position - point) for a point light (reference) or `direction` when
 * directional != 0. */
typedef struct {
    float position[3];
    float direction[3];
    int directional;
} cr_light;

/* Camera (src/Camera.cpp:59-60): V = lookAt(eye, target, up),
 * P = perspective(fovy, aspect, znear, zfar), glm's arithmetic: the
 * projection scales are tanf(fovy / 2) * aspect and tanf(fovy / 2), so any
 * fovy with a positive tangent of its half is accepted (the reference passes
 * fovy = 45, which glm takes as radians, and aspect = (float)(1280 / 720) = 1
 * after an integer division).  aspect 0 means width / height of the image.
 *
 * use_matrices = 1 takes the matrices ConeTraceShader uploads instead
 * (src/Shaders/ConeTraceShader.cpp:64-69), column-major as glm stores them:
 * V is used as given (its rotation rows orient the billboards, its third row
 * is the noise march's view ray) and P must be a symmetric perspective
 * projection (P[11] = -1, P[15] = 0, no shear or offset terms): the scales
 * are 1 / P[0] and 1 / P[5], znear = P[14] / (P[10] - 1) and
 * zfar = P[14] / (P[10] + 1); target, up, fovy, aspect, znear and zfar are
 * ignored.  eye stays the point the billboards are sorted from
 * (Camera::getPosition(), src/Shaders/ConeTraceShader.cpp:20). */
typedef struct {
    float eye[3];
    float target[3];
    float up[3];
    float fovy; /* radians */
    float znear;
    float zfar;
    float aspect;     /* 0: width / height */
    int use_matrices; /* 1: P and V below replace target, up, fovy, aspect, znear, zfar */
    float P[16];      /* column-major projection */
    float V[16];      /* column-major view */
} cr_camera;

/* Cone-trace parameters (src/Shaders/ConeTraceShader.hpp:28-32 defaults:
 * steps 16, angle 0.9, initial height 0.1, lod offset 0, down scaling 1). */
typedef struct {
    int steps;
    float cone_angle;     /* full aperture, radians */
    float initial_height; /* voxels */
    float lod_offset;
    float down_scaling;
    int cone_mode;        /* CR_CONE_* */
    float alpha_stop;     /* early-out threshold of the front-to-back modes */
    float sigma;          /* extinction per unit texture length (CR_CONE_FTB_DENSITY) */
    int composite;        /* CR_COMPOSITE_* */
    int quantize_rgba8;   /* GL_BLEND into an RGBA8 framebuffer: round after every blend */
    int compensated;      /* Kahan-compensated front-to-back accumulation */
    float clear_color[4]; /* glClearColor (src/main.cpp:112), GL_BLEND composite only */
} cr_cone_params;

/* Profiling results.  Level 1: counters and stage times of the last call.
 * Level 2: stage times summed over every call since the last cr_get_stats
 * (n_voxelize / n_cone calls), counters zero.  Stage times are hipEvent
 * spans on the stream each stage runs on: ms_clear covers the clear and the
 * pass-1 work list, ms_light_raster, ms_splat and ms_cone one kernel each,
 * ms_mip the level build, ms_setup / ms_sort / ms_bin the camera pass. */
typedef struct {
    int64_t light_fragments; /* covered pass-1 fragments */
    int64_t shell_texels;    /* position-map texels with a > 0 */
    int64_t fragments;       /* shaded camera fragments */
    int64_t samples;         /* cone samples taken */
    int64_t two_level;       /* samples that blended two mip levels */
    float ms_clear, ms_light_raster, ms_splat, ms_mip;  /* voxelize stages */
    float ms_setup, ms_sort, ms_bin, ms_cone;           /* cone-trace stages */
    int64_t n_voxelize, n_cone;                         /* calls summed (level 2) */
    int64_t chord_stores;    /* chord-fill voxel stores of pass 1 (rgba mode) */
    int64_t mlp_rows;        /* cones whose far field the MLP predicted (learned trace) */
    int64_t mlp_fallbacks;   /* of those, cones that fell back to the plain march */
    int64_t mlp_evals;       /* batched MLP evaluations (one per wave and billboard) */
} cr_stats;

/* ---- context ---------------------------------------------------------- */
/* context flags */
#define CR_CONTEXT_DIRECT_ADDRESSING 1u /* front-to-back cone traces compute every sample's texel addresses
                                           instead of reading per-level offset tables from LDS (same images) */
int cr_context_create(int device, unsigned flags, cr_context *out);
int cr_context_destroy(cr_context ctx);
const char *cr_last_error(cr_context ctx);
/* profiling 0: off.  1: hipEvents around every stage plus fragment/sample
 * counters, one stream synchronisation per call.  2: hipEvents only, no
 * synchronisation; stage times accumulate until cr_get_stats, which waits for
 * the recorded events and resets the sums. */
int cr_context_set_profiling(cr_context ctx, int profiling);
int cr_get_stats(cr_context ctx, cr_stats *out);
int cr_abi_version(void);

/* ---- volume (CloudVolume, src/CloudVolume.cpp:7-55) -------------------- */
/* bounds_lo/hi are (xBounds.x, yBounds.x, zBounds.x) / (.y ...), relative
 * to position, as in CloudVolume.  dim in [1, 1624] (texel indices inside a
 * level are 32-bit), levels in [1, log2(dim) + 1]. */
int cr_volume_create(cr_context ctx, int dim, const float bounds_lo[3], const float bounds_hi[3],
                     const float position[3], int levels, int format, int mode, cr_volume *out);
int cr_volume_destroy(cr_volume vol);
/* Density modulation by fractal value noise (rgba mode), octaves = 0 disables. */
int cr_volume_set_noise(cr_volume vol, int octaves, uint32_t seed, float base_freq, float lacunarity,
                        float gain);
/* Soft-particle noise march: the noise branch of the cone-trace fragment
 * shader (res/conetrace_frag.glsl:103-174, colour :194-196) with the
 * parameters of ConeTraceShader (src/Shaders/ConeTraceShader.hpp:14-25,
 * defaults in brackets) and its 32^3 RGBA8_SNORM noise map
 * (src/Shaders/ConeTraceShader.cpp:127-159).  The map's random alpha bytes
 * come from `seed` in place of rand(); normals follow the map's gradient as
 * the reference computes it.  offsets is the octaveOffsets uniform
 * (windVel * runTime, ConeTraceShader.cpp:54-61); octave i adds offsets[i]
 * to all three coordinates, 0 past the third. */
typedef struct {
    int enabled;         /* doNoiseSample */
    uint32_t seed;
    float step_size;     /* stepSize [0.01] */
    float opacity;       /* noiseOpacity [4] */
    int octaves;         /* numOctaves [4], 1..16 */
    float freq_step;     /* freqStep [3] */
    float pers_step;     /* persStep [0.5] */
    float adjust_size;   /* adjustSize [40] */
    int min_steps;       /* minNoiseSteps [2], >= 1 */
    int max_steps;       /* maxNoiseSteps [8], min_steps..64 */
    float min_color;     /* minNoiseColor [0.2] */
    float color_scale;   /* noiseColorScale [0.45] */
    float offsets[3];
} cr_noise_march;

/* Switch the noise march of cr_cone_trace on (nm->enabled != 0, the noise map
 * is built on the host and uploaded) or off (nm NULL or enabled 0).  With it
 * on, a camera fragment whose view ray misses its sphere is discarded and the
 * others blend (col * indirect, runningOpacity * alpha), clamped to [0, 1].
 * cr_cone_trace_learned and cr_volume_segments reject a volume with it on. */
int cr_volume_set_noise_march(cr_volume vol, const cr_noise_march *nm);
/* The noise map bytes of the last cr_volume_set_noise_march that switched it
 * on: 32^3 texels of (r, g, b, a), x fastest (bytes >= 131072). */
int cr_volume_read_noise_map(cr_volume vol, int8_t *host_dst, size_t bytes);
/* Sun sprite (SunShader::render, src/Shaders/SunShader.cpp:7-42,
 * res/sun_frag.glsl:14-28; defaults src/main.cpp:40-43 in brackets): a
 * camera-facing quad of half-size outer_radius at the light position, drawn
 * after the clear and before the clouds (src/main.cpp:112-116) with the GL
 * blend; inside inner_radius inner_color at alpha 1, then a ring blending to
 * outer_color with alpha 1 - scale, discarded where scale > 0.99.  GL-blend
 * composite only (cr_cone_trace refuses it with the front-to-back composite);
 * enabled 0 or NULL switches it off. */
typedef struct {
    int enabled;
    float inner_color[3]; /* [1, 1, 1] */
    float outer_color[3]; /* [1, 1, 0] */
    float inner_radius;   /* [1], >= 0 */
    float outer_radius;   /* [2], > inner_radius */
} cr_sun_sprite;
int cr_volume_set_sun_sprite(cr_volume vol, const cr_sun_sprite *sprite);
/* CloudVolume::uploadBillboards (src/CloudVolume.cpp:139-164): billboard
 * centres relative to position (xyz, 3*count floats) and scales; the radius
 * used is scale * fluffiness.  count = 0 is allowed (nothing is drawn). */
int cr_volume_set_billboards(cr_volume vol, const float *xyz, const float *scales, int count,
                             float fluffiness);

/* ---- the two render passes ------------------------------------------- */
/* VoxelizeShader::voxelize (src/Shaders/VoxelizeShader.cpp:18-118): clear,
 * pass 1 from the sun's orthographic view into a light_w x light_h position
 * map with a depth test, pass 2 splat of 9 voxels per texel, mip chain.
 * The volume holds two sets of occupancy words and mip levels: the clear and
 * both passes (on one stream owned by the volume) and the level build (on a
 * second) write the set that is not current, concurrently with work already
 * queued on `stream` (the previous cone trace reads the current set);
 * `stream` waits for the build, after which the new set is current. */
int cr_voxelize(cr_volume vol, const cr_light *light, int light_w, int light_h, void *stream);

/* ConeTraceShader::coneTrace (src/Shaders/ConeTraceShader.cpp:15-82):
 * sorts billboards far to near from the camera (CloudVolume::sortBoards),
 * shades every camera fragment with one cone toward the sun through the
 * mip chain and composites into d_image (W x H, row pitch in bytes).
 * The camera-side work (setup, sort, binning) runs on a stream owned by the
 * volume into one of two buffer pairs that consecutive traces alternate
 * between, concurrently with work already queued on `stream` (the previous
 * cone trace, a voxelization); the cone kernel itself runs on `stream` after
 * it.  d_image is written only by work on `stream`. */
int cr_cone_trace(cr_volume vol, const cr_light *light, const cr_camera *camera,
                  const cr_cone_params *params, void *d_image, int width, int height,
                  size_t pitch_bytes, int pixel_format, void *stream);

/* ---- readback (synchronise `stream`) ---------------------------------- */
/* Occupied level-0 voxels (lit shell and chord fill) and an
 * order-independent 64-bit hash: sum over occupied voxels of
 * splitmix64(linear_index * 4 + state), state 2 = lit shell, 1 = chord fill
 * only; linear_index = x + N*(y + N*z). */
int cr_volume_occupancy(cr_volume vol, uint64_t *count, uint64_t *hash, void *stream);
/* VoxelShader's activeVoxels (src/Shaders/VoxelShader.cpp:110-115): level-0
 * texels whose r is non-zero.  Those are the lit-shell voxels (chord-fill
 * voxels hold (0, 0, 0, 1) and noise scales only a), so in reference mode it
 * equals cr_volume_occupancy's count and in rgba mode it leaves the fill out. */
int cr_volume_active_voxels(cr_volume vol, uint64_t *count, void *stream);
/* Level `level` in x-fastest linear order and native element type
 * (u8 / float / binary16), channels 1 (R8) or 4. */
int cr_volume_read_level(cr_volume vol, int level, void *host_dst, size_t bytes, void *stream);
/* level-0 state bytes (0 empty, 1 fill, 2 shell), N^3, x-fastest */
int cr_volume_read_occupancy(cr_volume vol, uint8_t *host_dst, size_t bytes, void *stream);
/* pass-1 position map (RGBA32F, light_w x light_h) of the last cr_voxelize,
 * rebuilt from its depth test on request (CR_ESTATE if the billboards
 * changed since) */
int cr_volume_read_position_map(cr_volume vol, float *host_dst, size_t bytes, void *stream);
/* draw order of the last cr_cone_trace: order[r] = billboard drawn r-th (far to near) */
int cr_volume_read_draw_order(cr_volume vol, int32_t *host_dst, int count, void *stream);

/* ---- draw order as the reference keeps it ------------------------------ */
/* CR_ORDER_SORTED (default): cr_cone_trace sorts the billboards far to near
 * from the camera position first (CloudVolume::sortBoards) and draws them in
 * that order.  CR_ORDER_UPLOAD: cr_cone_trace draws them in upload order and
 * sorts nothing.  The reference does the latter: ConeTraceShader::coneTrace
 * sorts its CPU arrays in place (src/Shaders/ConeTraceShader.cpp:20,
 * src/CloudVolume.cpp:65-82) while its draw uses the buffers CloudVolume::
 * update uploaded at the start of the frame, so frame f draws (and voxelizes)
 * in the order sorted at frame f - 1, and each selection sort starts from
 * the previous order, which decides equal distances.  A caller reproduces
 * that with CR_ORDER_UPLOAD, cr_volume_sort_boards after the trace and the
 * permuted arrays uploaded by the next cr_volume_set_billboards. */
#define CR_ORDER_SORTED 0
#define CR_ORDER_UPLOAD 1
int cr_volume_set_draw_order(cr_volume vol, int order_mode);
/* CloudVolume::sortBoards(point) (src/CloudVolume.cpp:65-82) on the device:
 * host_order[r] = upload index of the billboard the selection sort puts at
 * position r (far to near from point; equal distances exactly as the
 * selection sort leaves them from the upload order).  count >= the number of
 * billboards.  Synchronous; touches no buffer a queued cone trace reads. */
int cr_volume_sort_boards(cr_volume vol, const float point[3], int32_t *host_order, int count, void *stream);

/* ---- learned far field ------------------------------------------------ */
/* No counterpart in the reference (it marches every cone to its end,
 * res/conetrace_frag.glsl:64-79).  A small MLP, 17 -> 64 -> 64 -> 64 -> 64 -> 5
 * with GELU (tanh form), BF16 weights and FP32 accumulation on the matrix
 * cores, predicts the far field of a front-to-back cone after its first k
 * samples.  Inputs (17): last sample position (2x - 1, texture space), cone
 * direction, sun direction, (density, light) at LODs 3..6 sampled along the
 * cone where its radius matches the LOD.  Outputs (5): far-field (r, g, b, a)
 * in units of the remaining transmittance, and a predicted error. */
#define CR_MLP_INPUTS 17
#define CR_MLP_OUTPUTS 5
#define CR_MLP_PARAMS 13957 /* W0[64][17] b0[64] W1[64][64] b1 W2 b2 W3 b3 W4[5][64] b4[5] */

int cr_mlp_create(cr_context ctx, cr_mlp *out); /* parameters zero */
int cr_mlp_destroy(cr_mlp mlp);
/* FP32 master parameters (host arrays of CR_MLP_PARAMS floats); setting
 * them resets the optimiser state.  Both synchronise the device. */
int cr_mlp_set_params(cr_mlp mlp, const float *host_params, int count);
int cr_mlp_get_params(cr_mlp mlp, float *host_params_out, int count);
/* d_y[n][5] = MLP(d_x[n][17]) (device arrays) */
int cr_mlp_forward(cr_mlp mlp, const float *d_x, int n, float *d_y, void *stream);
/* Training rows: d_x[n][17] features, d_t[n][5] targets (r, g, b, a,
 * remaining transmittance w).  Loss (mean over rows):
 *   w^2 sum_c (y_c - t_c)^2 + (y_4 - e)^2,  e = max_c |clamp(y_c, 0, 1) - t_c| (held constant).
 * cr_mlp_gradient: gradient (CR_MLP_PARAMS floats) and loss into host memory,
 * synchronous, parameters unchanged.  cr_mlp_train_step: one Adam step
 * (beta 0.9 / 0.999, eps 1e-8) with learning rate lr, asynchronous;
 * cr_mlp_last_loss reads the loss of the last step (synchronous). */
int cr_mlp_gradient(cr_mlp mlp, const float *d_x, const float *d_t, int n, float *host_grad, float *host_loss,
                    void *stream);
int cr_mlp_train_step(cr_mlp mlp, const float *d_x, const float *d_t, int n, float lr, void *stream);
int cr_mlp_last_loss(cr_mlp mlp, float *loss, void *stream);
/* No counterpart in the reference (it trains nothing); for callers that
 * cannot allocate: code under HIP graph capture, training loops that forbid
 * hipMalloc.  cr_mlp_train_step without allocation: the per-workgroup
 * gradient and loss partials of the step live in a fixed device array of the
 * library (CR_MLP_STATIC_ROWS / 64 workgroups x 13960 floats, about 14 MiB;
 * one copy per GPU, held until the process ends) instead of buffers that
 * cr_mlp_train_step grows with hipMalloc.  n must be in
 * [1, CR_MLP_STATIC_ROWS] (CR_EINVAL otherwise).  Same kernels and the same
 * result as cr_mlp_train_step.  Every call on one GPU shares that array
 * (every call writes all the partials it reads, so nothing needs clearing):
 * calls that overlap on one GPU -- on other streams, from other cr_mlp
 * objects or contexts -- race. */
#define CR_MLP_STATIC_ROWS 16384
int cr_mlp_train_step_static(cr_mlp mlp, const float *d_x, const float *d_t, int n, float lr, void *stream);

/* Training rows from the plain march: for start points d_p0[n][3] (texture
 * space) the front-to-back cone toward the directional light is marched as
 * cr_cone_trace marches it; cones still running after k samples get
 * d_valid = 1, their features, and targets ((c - c_k) / (1 - a_k),
 * (a - a_k) / (1 - a_k), 1 - a_k) from the rest of the march. */
int cr_volume_segments(cr_volume vol, const cr_light *light, const cr_cone_params *params, int k,
                       const float *d_p0, int n, float *d_features, float *d_targets, int32_t *d_valid,
                       void *stream);
/* cr_cone_trace with the far field of every cone still running after k
 * samples predicted by the MLP; a cone whose predicted error times its
 * remaining transmittance exceeds err_bound falls back to the plain march
 * (err_bound < 0: every cone falls back, giving cr_cone_trace's image).
 * Front-to-back cone and composite, RGBA32F / RGBA16F volumes. */
int cr_cone_trace_learned(cr_volume vol, const cr_light *light, const cr_camera *camera,
                          const cr_cone_params *params, cr_mlp mlp, int k, float err_bound, void *d_image,
                          int width, int height, size_t pitch_bytes, int pixel_format, void *stream);

/* ---- batched scenes ---------------------------------------------------- */
/* No counterpart in the reference (one scene per frame, src/main.cpp:98-124).
 * Many independent scenes rendered by one hipGraph: every scene's voxelize
 * and draw-order / binning passes, then a single cone-trace launch over the
 * camera tiles of all scenes, heaviest tiles first.  Volumes are created and
 * given billboards by the caller (distinct volumes, one context, RGBA32F or
 * RGBA16F, shared cone mode and composite); the graph is captured on the
 * first run and captured again when a configuration or a volume's billboards
 * or noise change.  Results equal per-scene cr_voxelize + cr_cone_trace. */
typedef struct {
    cr_volume volume;
    cr_light light;
    int light_w, light_h;
    cr_camera camera;
    cr_cone_params params;
    void *d_image;
    int width, height;
    size_t pitch_bytes;
    int pixel_format;
} cr_batch_item;

/* launch shape: camera tiles of tile_w x (64 / tile_w) pixels (one wave per
 * tile), waves_per_group waves per workgroup, and `groups` persistent
 * workgroups striding over the tiles (0: one workgroup per waves_per_group
 * tiles).  Default {8, 1, 0}. */
typedef struct {
    int tile_w;          /* 4, 8, 16 or 32 */
    int waves_per_group; /* 1, 2 or 4 */
    int groups;          /* 0 or the persistent grid size */
} cr_batch_config;

int cr_batch_create(cr_context ctx, const cr_batch_item *items, int count, cr_batch *out);
int cr_batch_destroy(cr_batch batch);
int cr_batch_set_config(cr_batch batch, const cr_batch_config *config);
int cr_batch_get_config(cr_batch batch, cr_batch_config *config);
/* Profiling level 1: eager run with counters; ms_setup = voxelize of all
 * scenes, ms_bin = camera setup + sort + binning of all scenes, ms_cone = the
 * batched cone trace.  Otherwise the graph is launched on `stream`. */
int cr_batch_run(cr_batch batch, void *stream);
/* Times every launch shape (tile_w 8/16/4 x waves 1/2/4 x one workgroup per
 * tile or 2048 persistent waves), `repeats` graph launches each, keeps the
 * fastest and returns it with its time per run. */
int cr_batch_autotune(cr_batch batch, int repeats, void *stream, cr_batch_config *best, float *best_ms);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDVOX_H_ */
