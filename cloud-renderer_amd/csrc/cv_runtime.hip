// Host runtime behind the C ABI (include/cloudvox.h).
//
// Replaces the GL state the reference's render passes drive:
//   VoxelizeShader::voxelize   src/Shaders/VoxelizeShader.cpp:18-134
//   ConeTraceShader::coneTrace src/Shaders/ConeTraceShader.cpp:15-98
//   CloudVolume (storage, billboards, sortBoards) src/CloudVolume.cpp:7-164
//   Sun::update                src/Sun.hpp:26-43
// Host-side matrix math follows glm 0.9.8 (lookAtRH, normalize =
// v * (1 / sqrt(dot)), dot = (x*x + y*y) + z*z) and is compiled with
// -ffp-contract=off so the view parameters are reproducible bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/cloudvox.h"
#include "cv_kernels.h"
#include "cv_mlp_launch.h"

using namespace cv;

struct cr_context_s {
    int device = 0;
    unsigned flags = 0;  // CR_CONTEXT_*
    float* static_scratch = nullptr;  // this device's copy of `scratch` (cr_mlp_train_step_static)
    // The context's three side streams, shared by its volumes and created
    // with the context.  The runtime backs streams with a few hardware queues
    // (four by default) and streams beyond that share one, which serialises
    // them; with the caller's stream these four need no more.  lside takes
    // the short latency-bound chains: the clear and the pass-1 work list
    // (built ahead of the raster) and the camera pass (setup, sort, binning);
    // vside the raster and the splat; bside the level build, which overlaps
    // the previous frame's cone trace on the caller's stream.
    hipStream_t lside = nullptr, vside = nullptr, bside = nullptr;
    hipEvent_t ev_side[3] = {nullptr, nullptr, nullptr};  // ordering points on them (cr_batch_run)
    std::string err;
    int profiling = 0;
    cr_stats stats{};  // level 1: last call; level 2: sums since the last cr_get_stats
    // event pool: level 1 reuses the first 10 events, level 2 takes 5 per call
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    struct Pending { int kind; size_t base; };  // kind 0 voxelize, 1 cone trace
    std::vector<Pending> pending;
};

// pass-1 work-list buffers of one voxelization
struct LightBufs {
    int4* lbox = nullptr;
    int* lcounts = nullptr;
    RasterItem* litems = nullptr;
    unsigned long long* keys = nullptr;
    size_t caps[4] = {0, 0, 0, 0};  // lbox, lcounts, litems, keys
};

// camera-pass buffers of one cone trace (billboard view data, draw order,
// tile masks, sort and binning scratch)
struct CamBufs {
    float4* sph_world = nullptr;
    float4* sph_view = nullptr;
    int4* sph_tiles = nullptr;
    unsigned long long* sort = nullptr;
    int32_t* order = nullptr;
    int* tie = nullptr;
    int* tiles = nullptr;
    unsigned long long* masks = nullptr;
    int sph_cap = 0;
    size_t sort_cap = 0, masks_cap = 0, tiles_cap = 0;
};

struct cr_volume_s {
    cr_context ctx = nullptr;
    int N = 0, L = 0, format = 0, mode = 0;
    // soft-particle noise march: parameters, the map's bytes and its decoded
    // texels on the device (null while the march is off)
    bool nm_on = false;
    cr_noise_march nm{};
    std::vector<int8_t> nm_bytes;
    float4* d_nmap = nullptr;
    // sun sprite drawn before the clouds (GL blend composite)
    bool sun_on = false;
    cr_sun_sprite sun{};
    float blo[3] = {}, bhi[3] = {}, pos[3] = {};
    VolumeDesc vd{};
    GridDesc g{};
    void* d_levels = nullptr;  // the current level set: the last build's, read by every reader of the levels
    size_t levels_bytes = 0;
    // two level sets: while the cone trace of frame f reads one, the build of
    // frame f + 1 writes the other (the second is allocated by the first
    // overlapped voxelization; without it the build runs on the caller's
    // stream into the one set).  ev_release[i] marks, on the caller's
    // stream, the end of the work enqueued while set i was current.
    void* d_lset[2] = {nullptr, nullptr};
    uint8_t* d_zflags[2] = {nullptr, nullptr};  // per set: regions whose slots hold zero (fused build)
    void* d_ntex = nullptr;  // density noise texture (vd.ntex), rebuilt by cr_volume_set_noise
    size_t ntex_cap = 0;
    size_t zflags_bytes = 0;
    int cur = 0;
    bool lset_failed = false;
    hipEvent_t ev_release[2] = {nullptr, nullptr};
    bool release_valid[2] = {false, false};
    unsigned long long* d_shell = nullptr;  // occupancy words of the current set (the last voxelization's)
    unsigned long long* d_fill = nullptr;
    // occupancy words per set, paired with the level sets: frame f + 1's
    // passes write one pair while frame f's build reads the other
    unsigned long long* d_shell_s[2] = {nullptr, nullptr};
    unsigned long long* d_fill_s[2] = {nullptr, nullptr};
    uint32_t nbricks0 = 0;
    int S = 0;
    float max_radius = 0.0f;
    float4* d_boards = nullptr;
    int boards_cap = 0;
    // pass 1
    int lw = 0, lh = 0;
    unsigned long long* d_keys = nullptr;
    float4* d_posmap = nullptr;
    size_t light_cap = 0;
    int4* d_lbox = nullptr;
    int* d_lcounts = nullptr;  // counts (S) followed by offsets (S + 1)
    RasterItem* d_litems = nullptr;
    size_t lbox_cap = 0, lcounts_cap = 0, litems_cap = 0;
    // consecutive overlapped voxelizations alternate between the work-list
    // and depth-key buffers above and light_spare (lset_i: which pair is
    // above); the clear and the work list run on the context's lside ahead of
    // the frame's raster, ev_items marks them done, ev_lfree[i] the end of
    // the last passes that read pair i
    LightBufs light_spare;
    int lset_i = 0;
    hipStream_t lside = nullptr;  // the context's, once used
    hipEvent_t ev_items = nullptr;
    hipEvent_t ev_lfree[2] = {nullptr, nullptr};
    bool lfree_valid[2] = {false, false};
    // camera pass
    float4* d_sph_world = nullptr;
    float4* d_sph_view = nullptr;
    int4* d_sph_tiles = nullptr;
    unsigned long long* d_sort = nullptr;
    int32_t* d_order = nullptr;
    int* d_tie = nullptr;
    int order_mode = CR_ORDER_SORTED;
    // cr_volume_sort_boards: keys and sort scratch, order, tie flag
    unsigned long long* d_xsort = nullptr;
    size_t xsort_cap = 0;
    int32_t* d_xorder = nullptr;
    size_t xorder_cap = 0;
    int* d_xtie = nullptr;
    size_t xtie_cap = 0;
    int* d_tiles = nullptr;  // binning / tile-queue scratch
    size_t tiles_cap = 0;
    int sph_cap = 0;
    size_t sort_cap = 0;
    unsigned long long* d_masks = nullptr;
    size_t masks_cap = 0;
    unsigned long long* d_counters = nullptr;
    uint8_t* d_scratch = nullptr;
    size_t scratch_cap = 0;
    bool voxelized = false;
    int sorted_count = -1;
    uint64_t gen = 0;  // bumped by every change of billboards / noise (invalidates captured batches)
    uint64_t vox_gen = 0;  // gen at the last voxelize
    LightView last_lv{};   // light view of the last voxelize
    size_t posmap_cap = 0;
    // camera-side stream of cr_cone_trace: cam setup, sort and binning run
    // there, overlapped with the caller's stream, which waits for them only
    // before the cone-trace kernel.  Consecutive traces alternate between the
    // buffers above and cam_spare (cset: which of the two pairs is above), so
    // a trace's camera pass overlaps the previous trace; ev_cfree[i] marks the
    // end of the last cone trace that read pair i, ev_csync orders the
    // camera stream after the caller's stream when cam_sync_needed.
    CamBufs cam_spare;
    int cset = 0;
    hipStream_t side = nullptr;  // the context's lside, once a trace has used it
    hipEvent_t ev_cam = nullptr, ev_csync = nullptr;
    hipEvent_t ev_cfree[2] = {nullptr, nullptr};
    bool cfree_valid[2] = {false, false};
    bool cam_sync_needed = true;
    // cr_voxelize: pass 1 and pass 2 run on the context's voxelization
    // stream (vside, once used), into the occupancy set that is not current,
    // overlapping the cone trace that reads the current set; the level build
    // follows on the context's bside after the splat (ev_splat), into the
    // level set that is not current, and the caller's stream waits for it
    // (ev_built).  ev_bfree[i]: the end of the last build that read occupancy
    // set i.  ev_sync orders the side streams after the caller's stream when
    // s_sync_needed (first call, after a serial call or a batch).
    hipStream_t vside = nullptr, bside = nullptr;
    hipEvent_t ev_splat = nullptr, ev_built = nullptr, ev_sync = nullptr;
    hipEvent_t ev_bfree[2] = {nullptr, nullptr};
    bool bfree_valid[2] = {false, false};
    bool s_sync_needed = true;
};

// far-field MLP: FP32 master parameters + Adam moments, bf16 packed copies for
// the matrix cores, per-workgroup gradient partials of the last training step
struct cr_mlp_s {
    cr_context ctx = nullptr;
    float* d_master = nullptr;
    float* d_m1 = nullptr;
    float* d_m2 = nullptr;
    uint16_t* d_fwd = nullptr;
    uint16_t* d_bwd = nullptr;
    float* d_bias = nullptr;
    float* d_grad = nullptr;
    float* d_loss = nullptr;
    float* d_partial = nullptr;
    float* d_loss_part = nullptr;
    size_t partial_cap = 0;  // workgroups
    size_t loss_cap = 0;
    int64_t step = 0;
};

static inline MlpDev mlp_dev(cr_mlp m) {
    MlpDev d;
    d.fwd = m->d_fwd;
    d.bwd = m->d_bwd;
    d.bias = m->d_bias;
    return d;
}

static thread_local std::string g_last_error;

namespace {

int fail(cr_context ctx, int code, const char* fmt, const char* a = nullptr) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), fmt, a ? a : "");
    if (ctx) ctx->err = buf;
    g_last_error = buf;
    return code;
}

#define CR_HIP(ctx, call)                                                                             \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail((ctx), CR_EDEVICE, #call ": %s", hipGetErrorString(e_));     \
    } while (0)

#define CR_HIP_ALLOC(ctx, call)                                                                       \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ == hipErrorOutOfMemory) return fail((ctx), CR_ENOMEM, #call ": out of device memory"); \
        if (e_ != hipSuccess) return fail((ctx), CR_EDEVICE, #call ": %s", hipGetErrorString(e_));     \
    } while (0)

// ---- glm 0.9.8 float math ----
struct F3 { float x, y, z; };
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline float dot3(F3 a, F3 b) {
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return (x + y) + z;
}
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }
inline F3 norm3(F3 a) { return a * (1.0f / std::sqrt(dot3(a, a))); }
inline F3 cross3(F3 a, F3 b) { return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

void look_at(F3 eye, F3 center, F3 up, float M[16]) {
    F3 f = norm3(center - eye);
    F3 s = norm3(cross3(f, up));
    F3 u = cross3(s, f);
    for (int i = 0; i < 16; ++i) M[i] = 0.0f;
    M[0] = s.x; M[4] = s.y; M[8] = s.z;
    M[1] = u.x; M[5] = u.y; M[9] = u.z;
    M[2] = -f.x; M[6] = -f.y; M[10] = -f.z;
    M[12] = -dot3(s, eye);
    M[13] = -dot3(u, eye);
    M[14] = dot3(f, eye);
    M[15] = 1.0f;
}

void basis_from_view(const float V[16], float S[3], float U[3], float B[3]) {
    S[0] = V[0]; S[1] = V[4]; S[2] = V[8];
    U[0] = V[1]; U[1] = V[5]; U[2] = V[9];
    F3 b = norm3(f3(V[2], V[6], V[10]));
    B[0] = b.x; B[1] = b.y; B[2] = b.z;
}

// Sun::update (src/Sun.hpp:26-43).  Up falls back to +z when the light looks
// along +-y (lookAt with up = +y is singular there).
void sun_update(const cr_volume_s* v, const float sun_pos[3], int W, int H, LightView* lv) {
    F3 mn = f3(v->blo[0], v->blo[1], v->blo[2]);
    F3 mx = f3(v->bhi[0], v->bhi[1], v->bhi[2]);
    F3 vpos = f3(v->pos[0], v->pos[1], v->pos[2]);
    F3 spos = f3(sun_pos[0], sun_pos[1], sun_pos[2]);
    F3 lookDir = norm3(vpos - spos);
    float ml = std::max(len3(mn), len3(mx));
    F3 lookPos = vpos - lookDir * ml;
    F3 up = f3(0.0f, 1.0f, 0.0f);
    if (std::fabs(lookDir.y) > 0.999f) up = f3(0.0f, 0.0f, 1.0f);
    look_at(lookPos, vpos, up, lv->V);
    float minmin = 2.0f * std::min(mn.x, std::min(mn.y, mn.z));
    float maxmax = 2.0f * std::max(mx.x, std::max(mx.y, mx.z));
    F3 nearPlane = lookPos + lookDir * 0.01f;
    F3 farPlane = lookPos + (lookDir * 2.0f) * ml;
    F3 dd = farPlane - nearPlane;
    float clip = len3(dd);
    basis_from_view(lv->V, lv->S, lv->U, lv->B);
    lv->np[0] = nearPlane.x; lv->np[1] = nearPlane.y; lv->np[2] = nearPlane.z;
    lv->clip = clip;
    // glm::ortho(minmin, maxmax, minmin, maxmax, 0.01, 0.01 + clip)
    lv->lsx = (maxmax - minmin) * 0.5f;
    lv->lox = (maxmax + minmin) * 0.5f;
    lv->lsy = (maxmax - minmin) * 0.5f;
    lv->loy = (maxmax + minmin) * 0.5f;
    lv->znear = 0.01f;
    lv->zfar = 0.01f + clip;
    lv->W = W;
    lv->H = H;
}

int level_dim(int N, int k) { return std::max(1, N >> k); }

template <class T>
int ensure(cr_context ctx, T** p, size_t need, size_t* cap) {
    if (need <= *cap && *p) return CR_OK;
    if (*p) {
        // work queued on the volume's streams may still read the old buffer
        // (growth is rare: wait for the device before freeing it)
        (void)hipDeviceSynchronize();
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipErrorOutOfMemory) { *p = nullptr; return fail(ctx, CR_ENOMEM, "hipMalloc: out of device memory"); }
    if (e != hipSuccess) { *p = nullptr; return fail(ctx, CR_EDEVICE, "hipMalloc: %s", hipGetErrorString(e)); }
    *cap = need;
    return CR_OK;
}

int set_device(cr_context ctx) {
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ctx, CR_EDEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    return CR_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0f;
    return ms;
}

// Accumulate the stage times of every pending level-2 call into ctx->stats.
int drain_pending(cr_context ctx) {
    for (const auto& p : ctx->pending) {
        hipEvent_t* e = &ctx->pool[p.base];
        hipError_t err = hipEventSynchronize(e[p.kind == 0 ? 4 : 5]);
        if (err != hipSuccess) return fail(ctx, CR_EDEVICE, "hipEventSynchronize: %s", hipGetErrorString(err));
        float t[4];
        for (int k = 0; k < 3; ++k) t[k] = elapsed(e[k], e[k + 1]);
        t[3] = p.kind == 0 ? elapsed(e[5], e[4]) : elapsed(e[4], e[5]);
        cr_stats& s = ctx->stats;
        if (p.kind == 0) {
            s.ms_clear += t[0]; s.ms_light_raster += t[1]; s.ms_splat += t[2]; s.ms_mip += t[3];
            ++s.n_voxelize;
        } else {
            s.ms_setup += t[0]; s.ms_sort += t[1]; s.ms_bin += t[2]; s.ms_cone += t[3];
            ++s.n_cone;
        }
    }
    ctx->pending.clear();
    ctx->pool_used = 0;
    return CR_OK;
}

// Stage events of one API call, six slots: voxelize brackets clear, pass 1
// and pass 2 with slots 0..3 (on the voxelization stream) and the level
// build with 5..4; a cone trace brackets setup, sort and binning with slots
// 0..3 (on the camera-side stream) and the cone kernel with 4..5.
struct StageEvents {
    cr_context ctx = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t* e = nullptr;
    size_t base = 0;
    int kind = 0;

    int begin(cr_context c, hipStream_t st, int k) {
        ctx = c; s = st; kind = k; e = nullptr;
        if (!c->profiling) return CR_OK;
        if (c->profiling == 1) {
            base = (size_t)(6 * k);
            if (c->pool.size() < 12) {
                while (c->pool.size() < 12) {
                    hipEvent_t ev;
                    hipError_t err = hipEventCreate(&ev);
                    if (err != hipSuccess) return fail(c, CR_EDEVICE, "hipEventCreate: %s", hipGetErrorString(err));
                    c->pool.push_back(ev);
                }
            }
        } else {
            if (c->pending.size() >= 1024) {
                int rc = drain_pending(c);
                if (rc) return rc;
            }
            base = c->pool_used;
            while (c->pool.size() < base + 6) {
                hipEvent_t ev;
                hipError_t err = hipEventCreate(&ev);
                if (err != hipSuccess) return fail(c, CR_EDEVICE, "hipEventCreate: %s", hipGetErrorString(err));
                c->pool.push_back(ev);
            }
            c->pool_used += 6;
        }
        e = &c->pool[base];
        return CR_OK;
    }
    void mark(int i) {
        if (e) (void)hipEventRecord(e[i], s);
    }
    void mark_on(int i, hipStream_t st) {
        if (e) (void)hipEventRecord(e[i], st);
    }
    void end() {
        if (e && ctx->profiling == 2) ctx->pending.push_back({kind, base});
    }
};

}  // namespace

extern "C" {

int cr_abi_version(void) { return CR_ABI_VERSION; }

const char* cr_last_error(cr_context ctx) {
    if (ctx) return ctx->err.c_str();
    return g_last_error.c_str();
}

// Gradient and loss partials of cr_mlp_train_step_static: one fixed array per
// GPU (the device copy of this file-scope variable).  Its host-side name is
// not a device address; host code uses the pointer hipGetSymbolAddress
// returns for the current device, looked up once per device.
namespace {
constexpr int kStaticGroups = (CR_MLP_STATIC_ROWS + 63) / 64;
constexpr size_t kStaticPartials = (size_t)kStaticGroups * kMlpParamsPad;
}  // namespace
static __device__ __align__(256) float scratch[kStaticPartials + kStaticGroups];

static std::mutex g_scratch_mu;
static std::vector<float*> g_scratch_dev;  // per device ordinal

// the current device's copy of `scratch` (device = the current device)
static int static_scratch_of(int device, float** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize((size_t)device + 1, nullptr);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, CR_EDEVICE, "hipGetSymbolAddress(scratch): %s",
                        e != hipSuccess ? hipGetErrorString(e) : "null address");
        g_scratch_dev[device] = (float*)p;
    }
    *out = g_scratch_dev[device];
    return CR_OK;
}

int cr_context_create(int device, unsigned flags, cr_context* out) {
    if (!out) return fail(nullptr, CR_EINVAL, "cr_context_create: out is NULL");
    if (flags & ~CR_CONTEXT_DIRECT_ADDRESSING) return fail(nullptr, CR_EINVAL, "cr_context_create: unknown flags");
    *out = nullptr;
    if (device < 0) return fail(nullptr, CR_EINVAL, "cr_context_create: negative device");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(nullptr, CR_EDEVICE, "cr_context_create: no HIP device (%s)",
                                               e != hipSuccess ? hipGetErrorString(e) : "count 0");
    if (device >= n) return fail(nullptr, CR_EINVAL, "cr_context_create: device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, CR_EDEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    float* sp = nullptr;
    int rc = static_scratch_of(device, &sp);
    if (rc) return rc;
    cr_context ctx = new (std::nothrow) cr_context_s();
    if (!ctx) return fail(nullptr, CR_ENOMEM, "cr_context_create: host allocation failed");
    ctx->device = device;
    ctx->flags = flags;
    ctx->static_scratch = sp;
    if ((e = hipStreamCreateWithFlags(&ctx->lside, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->vside, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->bside, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ctx->ev_side[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ctx->ev_side[1], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ctx->ev_side[2], hipEventDisableTiming)) != hipSuccess) {
        cr_context_destroy(ctx);
        return fail(nullptr, CR_EDEVICE, "cr_context_create: hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return CR_OK;
}

int cr_context_destroy(cr_context ctx) {
    if (!ctx) return CR_EINVAL;
    (void)hipSetDevice(ctx->device);
    for (hipStream_t st : {ctx->lside, ctx->vside, ctx->bside}) {
        if (!st) continue;
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (hipEvent_t e : ctx->ev_side)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->pool) (void)hipEventDestroy(e);
    delete ctx;
    return CR_OK;
}

int cr_context_set_profiling(cr_context ctx, int profiling) {
    if (!ctx) return fail(nullptr, CR_EINVAL, "cr_context_set_profiling: ctx is NULL");
    if (profiling < 0 || profiling > 2) return fail(ctx, CR_EINVAL, "cr_context_set_profiling: level must be 0, 1 or 2");
    int rc = set_device(ctx);
    if (rc) return rc;
    if (ctx->profiling == 2 && (rc = drain_pending(ctx))) return rc;
    ctx->profiling = profiling;
    ctx->stats = cr_stats{};
    ctx->pool_used = 0;
    return CR_OK;
}

int cr_get_stats(cr_context ctx, cr_stats* out) {
    if (!ctx || !out) return fail(ctx, CR_EINVAL, "cr_get_stats: NULL argument");
    if (ctx->profiling == 2) {
        int rc = set_device(ctx);
        if (rc) return rc;
        if ((rc = drain_pending(ctx))) return rc;
        *out = ctx->stats;
        ctx->stats = cr_stats{};
        return CR_OK;
    }
    *out = ctx->stats;
    return CR_OK;
}

int cr_volume_create(cr_context ctx, int dim, const float bounds_lo[3], const float bounds_hi[3],
                     const float position[3], int levels, int format, int mode, cr_volume* out) {
    if (!ctx || !out || !bounds_lo || !bounds_hi || !position)
        return fail(ctx, CR_EINVAL, "cr_volume_create: NULL argument");
    *out = nullptr;
    // texel indices inside a level are 32-bit: ceil(dim / 4)^3 * 64 < 2^32
    if (dim < 1 || dim > 1624) return fail(ctx, CR_EINVAL, "cr_volume_create: dim must be in [1, 1624]");
    int maxl = 1;
    while ((dim >> maxl) >= 1) ++maxl;  // floor(log2(dim)) + 1
    if (levels < 1 || levels > maxl || levels > kMaxLevels)
        return fail(ctx, CR_EINVAL, "cr_volume_create: levels must be in [1, log2(dim) + 1]");
    if (format < CR_FORMAT_R8 || format > CR_FORMAT_RGBA16F) return fail(ctx, CR_EINVAL, "cr_volume_create: bad format");
    if (mode != CR_MODE_REF && mode != CR_MODE_RGBA) return fail(ctx, CR_EINVAL, "cr_volume_create: bad mode");
    if (!finite3(bounds_lo) || !finite3(bounds_hi) || !finite3(position))
        return fail(ctx, CR_EINVAL, "cr_volume_create: non-finite bounds or position");
    for (int a = 0; a < 3; ++a)
        if (!(bounds_hi[a] > bounds_lo[a])) return fail(ctx, CR_EINVAL, "cr_volume_create: empty bounds");
    int rc = set_device(ctx);
    if (rc) return rc;
    cr_volume v = new (std::nothrow) cr_volume_s();
    if (!v) return fail(ctx, CR_ENOMEM, "cr_volume_create: host allocation failed");
    v->ctx = ctx;
    v->N = dim; v->L = levels; v->format = format; v->mode = mode;
    for (int a = 0; a < 3; ++a) { v->blo[a] = bounds_lo[a]; v->bhi[a] = bounds_hi[a]; v->pos[a] = position[a]; }
    // level layout
    uint64_t base = 0;
    for (int k = 0; k < levels; ++k) {
        uint32_t n = (uint32_t)level_dim(dim, k);
        uint32_t nb = (n + 3u) / 4u;
        v->vd.lv[k].n = n;
        v->vd.lv[k].nb = nb;
        v->vd.lv[k].base = base;
        v->vd.lv[k].fn = (float)n;
        base += (uint64_t)nb * nb * nb * 64u;
    }
    v->vd.L = levels;
    v->vd.format = format;
    v->vd.mode = mode;
    v->vd.noise_octaves = 0;
    v->vd.ntex = nullptr;
    v->levels_bytes = base * texel_bytes(format);
    // grid: world bounds = position + bounds (src/Shaders/VoxelizeShader.cpp:124-126);
    // voxelSize = (bounds.y - bounds.x) / dimension (src/CloudVolume.cpp:50-54)
    float vs[3];
    for (int a = 0; a < 3; ++a) {
        float wlo = position[a] + bounds_lo[a];
        float whi = position[a] + bounds_hi[a];
        v->g.lo[a] = wlo;
        v->g.range[a] = whi - wlo;
        v->g.vpos[a] = position[a];
        vs[a] = (bounds_hi[a] - bounds_lo[a]) / (float)dim;
    }
    v->g.stepSize = std::min(vs[0], std::min(vs[1], vs[2]));
    // exact reciprocals for power-of-two extents (x / 2^k == x * 2^-k in IEEE arithmetic)
    auto is_pow2 = [](float x) {  // 2^k with 2^-k a normal float
        int e;
        return x >= 0x1p-125f && x <= 0x1p125f && std::frexp(x, &e) == 0.5f;
    };
    v->g.pow2_range = is_pow2(v->g.range[0]) && is_pow2(v->g.range[1]) && is_pow2(v->g.range[2]) ? 1 : 0;
    for (int a = 0; a < 3; ++a) v->g.inv_range[a] = 1.0f / v->g.range[a];
    v->g.pow2_N = is_pow2((float)dim) ? 1 : 0;
    v->g.pow2_step = is_pow2(v->g.stepSize) ? 1 : 0;
    v->g.inv_step = 1.0f / v->g.stepSize;
    v->g.invN = 1.0f / (float)dim;
    v->g.N = dim;
    v->g.nb0 = v->vd.lv[0].nb;
    v->nbricks0 = v->g.nb0 * v->g.nb0 * v->g.nb0;
    hipError_t e = hipMalloc(&v->d_lset[0], v->levels_bytes);
    v->d_levels = v->d_lset[0];
    v->zflags_bytes = (build_zero_flags(v->g, v->vd) + 7) & ~(size_t)7;
    if (e == hipSuccess && v->zflags_bytes) e = hipMalloc((void**)&v->d_zflags[0], v->zflags_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&v->d_shell_s[0], (size_t)v->nbricks0 * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&v->d_fill_s[0], (size_t)v->nbricks0 * 8);
    v->d_shell = v->d_shell_s[0];
    v->d_fill = v->d_fill_s[0];
    if (e == hipSuccess) e = hipMalloc((void**)&v->d_counters, 16 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&v->d_tie, 64);
    if (e != hipSuccess) {
        cr_volume_destroy(v);
        return fail(ctx, e == hipErrorOutOfMemory ? CR_ENOMEM : CR_EDEVICE, "cr_volume_create: hipMalloc: %s",
                    hipGetErrorString(e));
    }
    // zero the volume so reads before the first voxelize are defined
    launch_fill_u64((uint64_t*)v->d_levels, v->levels_bytes / 8, 0, nullptr);
    if (v->d_zflags[0]) launch_fill_u64((uint64_t*)v->d_zflags[0], v->zflags_bytes / 8, 0x0101010101010101ull, nullptr);
    launch_fill_u64((uint64_t*)v->d_shell, v->nbricks0, 0, nullptr);
    launch_fill_u64((uint64_t*)v->d_fill, v->nbricks0, 0, nullptr);
    e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        cr_volume_destroy(v);
        return fail(ctx, CR_EDEVICE, "cr_volume_create: %s", hipGetErrorString(e));
    }
    *out = v;
    return CR_OK;
}

int cr_volume_destroy(cr_volume v) {
    if (!v) return CR_EINVAL;
    if (v->ctx) (void)hipSetDevice(v->ctx->device);
    for (hipStream_t st : {v->side, v->vside, v->bside, v->lside})  // the context's streams: drain, keep
        if (st) (void)hipStreamSynchronize(st);
    for (hipEvent_t e : {v->ev_cam, v->ev_csync, v->ev_cfree[0], v->ev_cfree[1], v->ev_splat, v->ev_built, v->ev_sync, v->ev_release[0],
                         v->ev_release[1], v->ev_bfree[0], v->ev_bfree[1], v->ev_items, v->ev_lfree[0],
                         v->ev_lfree[1]})
        if (e) (void)hipEventDestroy(e);
    void* ptrs[] = {v->d_lset[0], v->d_lset[1], v->d_zflags[0], v->d_zflags[1], v->d_shell_s[0], v->d_shell_s[1], v->d_fill_s[0], v->d_fill_s[1],    v->d_boards,    v->d_keys,      v->d_posmap,
                    v->d_lbox,   v->d_lcounts, v->d_litems,  v->d_sph_world, v->d_sph_view,  v->d_sph_tiles,
                    v->d_sort,   v->d_order,   v->d_tie,     v->d_masks,     v->d_counters,  v->d_scratch, v->d_tiles,
                    v->cam_spare.sph_world, v->cam_spare.sph_view, v->cam_spare.sph_tiles, v->cam_spare.sort,
                    v->cam_spare.order, v->cam_spare.tie, v->cam_spare.tiles, v->cam_spare.masks,
                    v->light_spare.lbox, v->light_spare.lcounts, v->light_spare.litems,
                    v->light_spare.keys, v->d_nmap, v->d_xsort, v->d_xorder, v->d_xtie, v->d_ntex};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete v;
    return CR_OK;
}

int cr_volume_set_noise(cr_volume v, int octaves, uint32_t seed, float base_freq, float lacunarity, float gain) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_set_noise: vol is NULL");
    if (octaves < 0 || octaves > 16) return fail(v->ctx, CR_EINVAL, "cr_volume_set_noise: octaves must be in [0, 16]");
    if (!std::isfinite(base_freq) || !std::isfinite(lacunarity) || !std::isfinite(gain))
        return fail(v->ctx, CR_EINVAL, "cr_volume_set_noise: non-finite parameter");
    cr_context ctx = v->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    // queued builds read the descriptor's texture: let them finish first
    CR_HIP(ctx, hipDeviceSynchronize());
    v->vd.noise_octaves = octaves;
    v->vd.noise_seed = seed;
    v->vd.noise_base_freq = base_freq;
    v->vd.noise_lacunarity = lacunarity;
    v->vd.noise_gain = gain;
    v->vd.ntex = nullptr;
    v->gen += 1;
    // The noise is a function of the voxel and these parameters only: the
    // build reads it from a texture computed here (2 or 4 bytes per texel)
    // instead of evaluating the octaves per frame.  Without memory for it the
    // build evaluates them.
    const size_t bytes = noise_tex_bytes(v->g, v->vd);
    if (bytes == 0) return CR_OK;
    if (bytes > v->ntex_cap) {
        if (v->d_ntex) (void)hipFree(v->d_ntex);
        v->d_ntex = nullptr;
        v->ntex_cap = 0;
        if (hipMalloc(&v->d_ntex, bytes) != hipSuccess) {
            (void)hipGetLastError();
            v->d_ntex = nullptr;
            return CR_OK;
        }
        v->ntex_cap = bytes;
    }
    launch_noise_tex(v->g, v->vd, v->d_ntex, nullptr);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipDeviceSynchronize());
    v->vd.ntex = v->d_ntex;
    return CR_OK;
}

// initNoiseMap(32) (src/Shaders/ConeTraceShader.cpp:127-159) with a seeded
// stand-in for rand(): m[4 * i + c] = (r, g, b, a) bytes, i = x + 32 * (y + 32 * z).
//   a = (char)(u * (128 - -128) + -128) (Util::genRandom(-128, 128),
//       src/Util.hpp:23-33), u = (hash(seed, x, y, z) >> 8) / 2^24 with the
//       lattice hash of the density noise;
//   rgb = (char)(normalize(g) * 128), g_x = d(x + 1) - d(x - 1) / 0.5 (the
//       division binds to the second term, as the reference writes it),
//       d = a / 128, neighbours wrapped (getIndex, :100-113).
// (char)f is the x86 conversion: truncate to int32 (NaN and out-of-range
// values give INT_MIN), keep the low byte.
static uint32_t nm_hash(uint32_t seed, int ix, int iy, int iz) {
    uint32_t h = seed;
    h ^= (uint32_t)ix * 0x8da6b343u;
    h ^= (uint32_t)iy * 0xd8163841u;
    h ^= (uint32_t)iz * 0xcb1ab31fu;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
static int8_t nm_char(float v) {
    const int32_t i = (v == v && std::fabs(v) < 2147483648.0f) ? (int32_t)v : INT32_MIN;
    return (int8_t)(uint8_t)((uint32_t)i & 0xffu);
}
static void noise_map_build(uint32_t seed, std::vector<int8_t>& m) {
    constexpr int D = 32;
    m.assign((size_t)D * D * D * 4, 0);
    auto at = [](int x, int y, int z) {
        x = (x + D) % D; y = (y + D) % D; z = (z + D) % D;
        return x + D * (y + D * z);
    };
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) {
                const float u = (float)(nm_hash(seed, x, y, z) >> 8) * (1.0f / 16777216.0f);
                m[4 * at(x, y, z) + 3] = nm_char(u * (128.0f - -128.0f) + -128.0f);
            }
    auto d = [&](int i) { return (float)m[4 * i + 3] / 128.0f; };
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) {
                const F3 g = f3(d(at(x + 1, y, z)) - d(at(x - 1, y, z)) / 0.5f,
                                d(at(x, y + 1, z)) - d(at(x, y - 1, z)) / 0.5f,
                                d(at(x, y, z + 1)) - d(at(x, y, z - 1)) / 0.5f);
                const F3 n = norm3(g);
                const int i = at(x, y, z);
                m[4 * i + 0] = nm_char(n.x * 128.0f);
                m[4 * i + 1] = nm_char(n.y * 128.0f);
                m[4 * i + 2] = nm_char(n.z * 128.0f);
            }
}

int cr_volume_set_noise_march(cr_volume v, const cr_noise_march* nm) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_set_noise_march: vol is NULL");
    cr_context ctx = v->ctx;
    if (!nm || !nm->enabled) {
        v->nm_on = false;
        return CR_OK;
    }
    if (nm->octaves < 1 || nm->octaves > 16)
        return fail(ctx, CR_EINVAL, "cr_volume_set_noise_march: octaves must be in [1, 16]");
    // iSteps = min(len / stepSize, max - min) + min and the march divides by
    // iSteps - 1 (res/conetrace_frag.glsl:154-158): min = max = 1 would divide
    // by zero, so max_steps >= 2 (then iSteps - 1 > 0 for every kept fragment)
    if (nm->min_steps < 1 || nm->max_steps < nm->min_steps || nm->max_steps > 64 || nm->max_steps < 2)
        return fail(ctx, CR_EINVAL, "cr_volume_set_noise_march: need 1 <= min_steps <= max_steps <= 64, max_steps >= 2");
    const float fs[] = {nm->step_size, nm->opacity, nm->freq_step, nm->pers_step, nm->adjust_size, nm->min_color,
                        nm->color_scale, nm->offsets[0], nm->offsets[1], nm->offsets[2]};
    for (float f : fs)
        if (!std::isfinite(f)) return fail(ctx, CR_EINVAL, "cr_volume_set_noise_march: non-finite parameter");
    if (!(nm->step_size > 0.0f) || nm->adjust_size == 0.0f)
        return fail(ctx, CR_EINVAL, "cr_volume_set_noise_march: step_size must be > 0 and adjust_size != 0");
    int rc = set_device(ctx);
    if (rc) return rc;
    if (!v->d_nmap || !v->nm_on || v->nm.seed != nm->seed || v->nm_bytes.empty()) {
        std::vector<int8_t> m;
        noise_map_build(nm->seed, m);
        // decoded texels: RGBA8_SNORM -> max(c / 127, -1) (GL 4.4 eq. 2.2)
        std::vector<float4> t(m.size() / 4);
        for (size_t i = 0; i < t.size(); ++i)
            t[i] = make_float4(std::max((float)m[4 * i] / 127.0f, -1.0f), std::max((float)m[4 * i + 1] / 127.0f, -1.0f),
                               std::max((float)m[4 * i + 2] / 127.0f, -1.0f), std::max((float)m[4 * i + 3] / 127.0f, -1.0f));
        // cone traces still queued may read the old map
        CR_HIP(ctx, hipDeviceSynchronize());
        if (!v->d_nmap) CR_HIP(ctx, hipMalloc((void**)&v->d_nmap, t.size() * sizeof(float4)));
        CR_HIP(ctx, hipMemcpy(v->d_nmap, t.data(), t.size() * sizeof(float4), hipMemcpyHostToDevice));
        v->nm_bytes.swap(m);
    }
    v->nm = *nm;
    v->nm_on = true;
    return CR_OK;
}

int cr_volume_set_sun_sprite(cr_volume v, const cr_sun_sprite* sp) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_set_sun_sprite: vol is NULL");
    if (!sp || !sp->enabled) {
        v->sun_on = false;
        return CR_OK;
    }
    if (!finite3(sp->inner_color) || !finite3(sp->outer_color) || !std::isfinite(sp->inner_radius) ||
        !std::isfinite(sp->outer_radius) || !(sp->inner_radius >= 0.0f) || !(sp->outer_radius > sp->inner_radius))
        return fail(v->ctx, CR_EINVAL, "cr_volume_set_sun_sprite: need finite colours and 0 <= inner < outer radius");
    v->sun = *sp;
    v->sun_on = true;
    return CR_OK;
}

int cr_volume_read_noise_map(cr_volume v, int8_t* host_dst, size_t bytes) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_read_noise_map: vol is NULL");
    if (!host_dst) return fail(v->ctx, CR_EINVAL, "cr_volume_read_noise_map: NULL destination");
    if (v->nm_bytes.empty()) return fail(v->ctx, CR_ESTATE, "cr_volume_read_noise_map: the noise march was never on");
    if (bytes < v->nm_bytes.size()) return fail(v->ctx, CR_EINVAL, "cr_volume_read_noise_map: need 131072 bytes");
    std::memcpy(host_dst, v->nm_bytes.data(), v->nm_bytes.size());
    return CR_OK;
}

int cr_volume_set_billboards(cr_volume v, const float* xyz, const float* scales, int count, float fluffiness) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_set_billboards: vol is NULL");
    cr_context ctx = v->ctx;
    if (count < 0) return fail(ctx, CR_EINVAL, "cr_volume_set_billboards: negative count");
    if (count > 0 && (!xyz || !scales)) return fail(ctx, CR_EINVAL, "cr_volume_set_billboards: NULL array");
    if (!std::isfinite(fluffiness)) return fail(ctx, CR_EINVAL, "cr_volume_set_billboards: non-finite fluffiness");
    int rc = set_device(ctx);
    if (rc) return rc;
    std::vector<float4> h((size_t)std::max(count, 1));
    float maxr = 0.0f;
    for (int j = 0; j < count; ++j) {
        // uploadBillboards (src/CloudVolume.cpp:153-160): scale * fluffiness
        float r = fluffiness != 1.0f ? scales[j] * fluffiness : scales[j];
        if (!finite3(xyz + 3 * j) || !std::isfinite(r))
            return fail(ctx, CR_EINVAL, "cr_volume_set_billboards: non-finite position or scale");
        h[j] = make_float4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], r);
        maxr = std::max(maxr, r);
    }
    // pending camera or voxelization passes read the boards
    if (v->side) CR_HIP(ctx, hipStreamSynchronize(v->side));
    if (v->vside) CR_HIP(ctx, hipStreamSynchronize(v->vside));
    if (v->bside) CR_HIP(ctx, hipStreamSynchronize(v->bside));
    if (v->lside) CR_HIP(ctx, hipStreamSynchronize(v->lside));
    size_t cap = (size_t)v->boards_cap;
    rc = ensure(ctx, &v->d_boards, (size_t)std::max(count, 1), &cap);
    if (rc) return rc;
    v->boards_cap = (int)cap;
    if (count > 0) CR_HIP(ctx, hipMemcpy(v->d_boards, h.data(), (size_t)count * sizeof(float4), hipMemcpyHostToDevice));
    v->S = count;
    v->max_radius = maxr;
    v->sorted_count = -1;
    v->gen += 1;
    return CR_OK;
}

// pass-1/pass-2/mip launches of one voxelize call, prepared on the host
struct VoxPlan {
    LightView lv;
    LightScratch ls;
    size_t npx;
    size_t nitems;  // work-list capacity
};

static int voxelize_prepare(cr_volume v, const cr_light* light, int light_w, int light_h, VoxPlan& plan) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_voxelize: vol is NULL");
    cr_context ctx = v->ctx;
    if (!light) return fail(ctx, CR_EINVAL, "cr_voxelize: light is NULL");
    if (light_w <= 0 || light_h <= 0 || light_w > 16384 || light_h > 16384)
        return fail(ctx, CR_EINVAL, "cr_voxelize: light map size must be in [1, 16384]");
    if (!finite3(light->position)) return fail(ctx, CR_EINVAL, "cr_voxelize: non-finite light position");
    {
        float d0 = v->pos[0] - light->position[0], d1 = v->pos[1] - light->position[1],
              d2 = v->pos[2] - light->position[2];
        if (d0 == 0.0f && d1 == 0.0f && d2 == 0.0f)
            return fail(ctx, CR_EINVAL, "cr_voxelize: light position coincides with the volume position");
    }
    // the chord fill walks 2 * radius in stepSize increments with a float
    // counter; bound it so the walk always advances and terminates
    if (v->mode == CR_MODE_RGBA && (2.0f * v->max_radius) / v->g.stepSize > 1048576.0f)
        return fail(ctx, CR_EUNSUPPORTED, "cr_voxelize: billboard radius exceeds 2^19 voxels (chord fill)");
    int rc = set_device(ctx);
    if (rc) return rc;
    sun_update(v, light->position, light_w, light_h, &plan.lv);
    size_t npx = (size_t)light_w * light_h;
    plan.npx = npx;
    v->lw = light_w;
    v->lh = light_h;
    v->last_lv = plan.lv;  // the position map is rebuilt from the keys on request
    v->vox_gen = v->gen;
    plan.ls = LightScratch{};
    plan.nitems = v->S > 0 ? light_items_bound(v->S, v->max_radius, plan.lv, v->g, v->mode == CR_MODE_RGBA ? 1 : 0) : 0;
    return CR_OK;
}

// the volume's current work-list buffers, grown to the plan
static int light_scratch(cr_volume v, VoxPlan& plan) {
    cr_context ctx = v->ctx;
    plan.ls = LightScratch{};
    int rc;
    if ((rc = ensure(ctx, &v->d_keys, plan.npx, &v->light_cap))) return rc;
    if (v->S <= 0) return CR_OK;
    if ((rc = ensure(ctx, &v->d_lbox, (size_t)v->S, &v->lbox_cap))) return rc;
    if ((rc = ensure(ctx, &v->d_lcounts, 2 * (size_t)v->S + 1, &v->lcounts_cap))) return rc;
    if ((rc = ensure(ctx, &v->d_litems, plan.nitems, &v->litems_cap))) return rc;
    plan.ls.lbox = v->d_lbox;
    plan.ls.counts = v->d_lcounts;
    plan.ls.offsets = v->d_lcounts + v->S;
    plan.ls.items = v->d_litems;
    plan.ls.items_cap = plan.nitems;
    return CR_OK;
}

// exchange the volume's work-list buffers with the spare pair
static void swap_light(cr_volume v) {
    LightBufs& o = v->light_spare;
    std::swap(v->d_lbox, o.lbox);
    std::swap(v->d_lcounts, o.lcounts);
    std::swap(v->d_litems, o.litems);
    std::swap(v->lbox_cap, o.caps[0]);
    std::swap(v->lcounts_cap, o.caps[1]);
    std::swap(v->litems_cap, o.caps[2]);
    std::swap(v->d_keys, o.keys);
    std::swap(v->light_cap, o.caps[3]);
    v->lset_i ^= 1;
}

// kernels only (no allocation, no host synchronisation): safe under graph capture
// The clear and the work list on `ls` (after `lwaits` when given), passes 1
// and 2 into occupancy set `set` on `vs`, the level build into level set
// `set` on `bs` (after them when the streams differ, and after `wait` when
// given).
static void voxelize_enqueue(cr_volume v, const VoxPlan& plan, unsigned long long* counters, StageEvents* ev,
                             hipStream_t ls, const hipEvent_t* lwaits, hipStream_t vs, hipStream_t bs, int set,
                             hipEvent_t wait) {
    unsigned long long* shell = v->d_shell_s[set];
    unsigned long long* fill = v->d_fill_s[set];
    if (lwaits)
        for (int i = 0; i < 2; ++i)
            if (lwaits[i]) (void)hipStreamWaitEvent(ls, lwaits[i], 0);
    if (ev) ev->mark_on(0, ls);
    // clearGPU + clearPositionMap (src/CloudVolume.cpp:96-100, VoxelizeShader.cpp:185-187), one launch
    const FillSeg segs[4] = {{(uint64_t*)v->d_keys, (uint64_t)plan.npx, ~0ull},
                             {(uint64_t*)shell, (uint64_t)v->nbricks0, 0},
                             {(uint64_t*)fill, (uint64_t)v->nbricks0, 0},
                             {(uint64_t*)counters, counters ? 9u : 0u, 0}};
    launch_fill_multi(segs, 4, ls);
    launch_light_items(v->g, plan.lv, v->mode, v->d_boards, v->S, plan.ls, ls);
    if (ls != vs) {
        (void)hipEventRecord(v->ev_items, ls);
        (void)hipStreamWaitEvent(vs, v->ev_items, 0);
    }
    // stage spans: [0, 1] clear + pass-1 work list, [1, 2] the raster kernel
    // alone, [2, 3] the splat kernel alone, [5, 4] the level build
    if (ev) ev->mark_on(1, vs);
    launch_light_raster(v->g, plan.lv, v->mode, v->d_boards, v->S, plan.ls, v->d_keys, fill, counters, vs);
    if (ev) ev->mark_on(2, vs);
    launch_light_splat(v->g, plan.lv, v->d_boards, v->d_keys, nullptr, shell, counters, vs);
    if (ls != vs) (void)hipEventRecord(v->ev_lfree[v->lset_i], vs);
    if (ev) ev->mark_on(3, vs);
    if (vs != bs) {
        (void)hipEventRecord(v->ev_splat, vs);
        (void)hipStreamWaitEvent(bs, v->ev_splat, 0);
    }
    if (wait) (void)hipStreamWaitEvent(bs, wait, 0);
    if (ev) ev->mark_on(5, bs);
    launch_build_levels(v->g, v->vd, shell, v->mode == CR_MODE_RGBA ? fill : nullptr, v->d_lset[set],
                        v->d_zflags[set], bs);
    if (ev) ev->mark_on(4, bs);
}

int cr_voxelize(cr_volume v, const cr_light* light, int light_w, int light_h, void* stream) {
    VoxPlan plan;
    int rc = voxelize_prepare(v, light, light_w, light_h, plan);
    if (rc) return rc;
    cr_context ctx = v->ctx;
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* counters = ctx->profiling == 1 ? v->d_counters : nullptr;
    // Clear and work list (on the context's work-list stream), pass 1 and
    // pass 2 (on its voxelization stream) write only the keys, the work list
    // and the occupancy set that is not current, so they overlap the cone
    // trace that reads the current sets; they wait for the end of the build
    // that last read their occupancy set.  The build (on the build stream)
    // follows the splat and the end of the work enqueued on the caller's
    // stream while its level set was current, and the caller's stream waits
    // for the build.  Without room for second sets, or in counter runs,
    // everything runs on the caller's stream.
    const bool overlap = ctx->profiling != 1;
    if (overlap) {
        if (!v->vside) {
            v->vside = ctx->vside;
            v->bside = ctx->bside;
            v->lside = ctx->lside;
            for (hipEvent_t* e : {&v->ev_splat, &v->ev_built, &v->ev_sync, &v->ev_release[0], &v->ev_release[1],
                                  &v->ev_bfree[0], &v->ev_bfree[1], &v->ev_items, &v->ev_lfree[0], &v->ev_lfree[1]})
                CR_HIP(ctx, hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
        if (!v->d_lset[1] && !v->lset_failed) {
            void* p = nullptr;
            uint8_t* z = nullptr;
            unsigned long long *sh = nullptr, *fl = nullptr;
            const size_t wb = (size_t)v->nbricks0 * 8;
            if (hipMalloc(&p, v->levels_bytes) != hipSuccess ||
                (v->zflags_bytes && hipMalloc((void**)&z, v->zflags_bytes) != hipSuccess) ||
                hipMalloc((void**)&sh, wb) != hipSuccess || hipMalloc((void**)&fl, wb) != hipSuccess) {
                (void)hipGetLastError();  // no room for second sets: one set on the caller's stream
                for (void* q : {(void*)p, (void*)z, (void*)sh, (void*)fl})
                    if (q) (void)hipFree(q);
                v->lset_failed = true;
            } else {
                v->d_lset[1] = p;
                v->d_zflags[1] = z;
                v->d_shell_s[1] = sh;
                v->d_fill_s[1] = fl;
                launch_fill_u64((uint64_t*)p, v->levels_bytes / 8, 0, s);
                if (z) launch_fill_u64((uint64_t*)z, v->zflags_bytes / 8, 0x0101010101010101ull, s);
                v->s_sync_needed = true;  // the streams wait for these fills
            }
        }
    }
    const bool two = overlap && v->d_lset[1] != nullptr;
    const int nxt = two ? v->cur ^ 1 : v->cur;
    hipStream_t vs = two ? v->vside : s, bs = two ? v->bside : s, ls = two ? v->lside : s;
    // the work list goes into the pair the previous voxelization did not use
    if (two) swap_light(v);
    if ((rc = light_scratch(v, plan))) {
        if (two) swap_light(v);
        return rc;
    }
    if (two) {
        if (v->s_sync_needed) {  // after all prior work on s
            CR_HIP(ctx, hipEventRecord(v->ev_sync, s));
            for (hipStream_t st : {vs, bs, ls}) CR_HIP(ctx, hipStreamWaitEvent(st, v->ev_sync, 0));
            v->s_sync_needed = false;
        }
    }
    // the clear waits for the passes that last read its key pair and the
    // build that last read its occupancy set
    const hipEvent_t lwaits[2] = {two && v->lfree_valid[v->lset_i] ? v->ev_lfree[v->lset_i] : nullptr,
                                  two && v->bfree_valid[nxt] ? v->ev_bfree[nxt] : nullptr};
    StageEvents ev;
    if ((rc = ev.begin(ctx, vs, 0))) return rc;
    voxelize_enqueue(v, plan, counters, &ev, ls, lwaits, vs, bs, nxt,
                     two && v->release_valid[nxt] ? v->ev_release[nxt] : nullptr);
    ev.end();
    if (two) v->lfree_valid[v->lset_i] = true;
    if (two) {
        CR_HIP(ctx, hipEventRecord(v->ev_bfree[nxt], bs));
        v->bfree_valid[nxt] = true;
        // what was enqueued on s so far may read the current level set
        CR_HIP(ctx, hipEventRecord(v->ev_release[v->cur], s));
        v->release_valid[v->cur] = true;
        CR_HIP(ctx, hipEventRecord(v->ev_built, bs));
        CR_HIP(ctx, hipStreamWaitEvent(s, v->ev_built, 0));
        v->cur = nxt;
        v->d_levels = v->d_lset[nxt];
        v->d_shell = v->d_shell_s[nxt];
        v->d_fill = v->d_fill_s[nxt];
    } else {
        v->s_sync_needed = true;
    }
    CR_HIP(ctx, hipGetLastError());
    if (ctx->profiling == 1) {
        CR_HIP(ctx, hipStreamSynchronize(s));
        unsigned long long c[8];
        CR_HIP(ctx, hipMemcpy(c, v->d_counters, sizeof(c), hipMemcpyDeviceToHost));
        ctx->stats.light_fragments = (int64_t)c[0];
        ctx->stats.chord_stores = (int64_t)c[1];
        ctx->stats.shell_texels = (int64_t)c[2];
        ctx->stats.ms_clear = elapsed(ev.e[0], ev.e[1]);
        ctx->stats.ms_light_raster = elapsed(ev.e[1], ev.e[2]);
        ctx->stats.ms_splat = elapsed(ev.e[2], ev.e[3]);
        ctx->stats.ms_mip = elapsed(ev.e[5], ev.e[4]);
        ctx->stats.n_voxelize = 1;
    }
    v->voxelized = true;
    return CR_OK;
}

// far-field options of cr_cone_trace_learned (null for the plain trace)
struct LearnedOpts {
    cr_mlp mlp;
    int k;
    float err_bound;
};

static void fill_cone_desc(cr_volume v, const cr_light* light, const cr_cone_params* p, ConeDesc& cd) {
    cd.steps = p->steps;
    float tanHalf = std::tan(p->cone_angle / 2.0f);
    cd.tanHalf = tanHalf;
    float h = p->initial_height;
    for (int i = 0; i < p->steps; ++i) {
        float r = h * tanHalf;
        float lod = std::log2(std::max(1.0f, 2.0f * r));
        cd.h[i] = h;
        // lambda -> level / fraction, in FP32 as the samplers did per sample
        const float lam = lod + p->lod_offset;
        int lv = 0;
        float f = 0.0f;
        if (lam <= 0.0f) {
        } else if (lam >= (float)(v->L - 1)) {
            lv = v->L - 1;
        } else {
            const float fl = std::floor(lam);
            lv = (int)fl;
            f = lam - fl;
        }
        cd.llev4[i >> 2] = (cd.llev4[i >> 2] & ~(0xffu << (8 * (i & 3)))) | ((uint32_t)lv << (8 * (i & 3)));
        cd.lfrac[i] = f;
        h = h + r;
    }
    cd.den = (float)p->steps * p->down_scaling;
    cd.alpha_stop = p->alpha_stop;
    cd.kscale = (p->sigma * (float)1.4426950408889634) / (float)v->N;
    cd.cone_mode = p->cone_mode;
    cd.count = 1;
    cd.composite = p->composite;
    cd.quantize = (p->composite == CR_COMPOSITE_GL_BLEND && p->quantize_rgba8) ? 1 : 0;
    cd.compensated = p->compensated ? 1 : 0;
    for (int q = 0; q < 4; ++q) cd.clear[q] = p->clear_color[q];
    for (int a = 0; a < 3; ++a) cd.lpos[a] = light->position[a];
    cd.directional = light->directional ? 1 : 0;
    F3 sun;
    if (cd.directional) {
        F3 ld = norm3(f3(light->direction[0], light->direction[1], light->direction[2]));
        float fN = (float)v->N;
        cd.ldir[0] = ld.x / fN;
        cd.ldir[1] = ld.y / fN;
        cd.ldir[2] = ld.z / fN;
        sun = ld;
    } else {
        F3 centre = f3(v->g.lo[0] + v->g.range[0] * 0.5f, v->g.lo[1] + v->g.range[1] * 0.5f,
                       v->g.lo[2] + v->g.range[2] * 0.5f);
        sun = norm3(f3(light->position[0], light->position[1], light->position[2]) - centre);
    }
    cd.sun[0] = sun.x; cd.sun[1] = sun.y; cd.sun[2] = sun.z;
    // far-field feature heights: the cone radius is 2^(lod-1) voxels (lambda =
    // lod) at height 2^(lod-1) / tan(angle / 2); LODs 3..6
    for (int j = 0; j < 4; ++j) {
        int lod = 3 + j;
        cd.feat_h[j] = (float)(1 << (lod - 1)) / tanHalf;
        cd.feat_lvl[j] = std::min(lod, v->L - 1);
    }
    cd.learned_k = 0;
    cd.err_bound = 0.0f;
}

static int cone_params_check(cr_context ctx, cr_volume v, const cr_cone_params* p, const char* who) {
    if (p->steps < 0 || p->steps > kMaxSteps) return fail(ctx, CR_EUNSUPPORTED, "%s: steps must be in [0, 256]", who);
    if (p->cone_mode < CR_CONE_WEIGHTED || p->cone_mode > CR_CONE_FTB_DENSITY) return fail(ctx, CR_EINVAL, "%s: bad cone mode", who);
    if (p->composite != CR_COMPOSITE_GL_BLEND && p->composite != CR_COMPOSITE_FTB) return fail(ctx, CR_EINVAL, "%s: bad composite", who);
    if (!(p->cone_angle > 0.0f && p->cone_angle < 3.14159265f)) return fail(ctx, CR_EINVAL, "%s: cone angle must be in (0, pi)", who);
    if (!(p->initial_height > 0.0f) || !std::isfinite(p->initial_height)) return fail(ctx, CR_EINVAL, "%s: initial height must be > 0", who);
    if (!std::isfinite(p->lod_offset) || !(p->down_scaling != 0.0f) || !std::isfinite(p->down_scaling))
        return fail(ctx, CR_EINVAL, "%s: bad lod offset / down scaling", who);
    if (!std::isfinite(p->sigma) || p->sigma < 0.0f) return fail(ctx, CR_EINVAL, "%s: sigma must be finite and >= 0", who);
    if (v->format == CR_FORMAT_R8 && p->cone_mode != CR_CONE_WEIGHTED)
        return fail(ctx, CR_EUNSUPPORTED, "%s: R8 volumes hold one channel; use the weighted cone", who);
    return CR_OK;
}

// descriptors and buffers of one cone-trace call, prepared on the host
struct ConePlan {
    CameraView cv;
    ConeDesc cd;
    int S;
    int words;
    size_t ntiles;
    size_t nmask;
};

// Camera (src/Camera.cpp:59-60) as ConeTraceShader uploads it
// (src/Shaders/ConeTraceShader.cpp:64-69): V = lookAt(eye, target, up) and
// glm::perspective's scales tanf(fovy / 2) * aspect, tanf(fovy / 2), or V and
// P given as matrices (scales 1 / P[0], 1 / P[5]; near and far from P[10],
// P[14]).  The billboard basis is V's rotation rows; eye is the sort point.
static int camera_view(cr_context ctx, const cr_camera* cam, int W, int H, CameraView& cv) {
    if (!finite3(cam->eye)) return fail(ctx, CR_EINVAL, "cr_cone_trace: non-finite camera");
    if (cam->use_matrices) {
        const float* P = cam->P;
        for (int i = 0; i < 16; ++i)
            if (!std::isfinite(P[i]) || !std::isfinite(cam->V[i]))
                return fail(ctx, CR_EINVAL, "cr_cone_trace: non-finite camera matrix");
        static const int zero[] = {1, 2, 3, 4, 6, 7, 8, 9, 12, 13, 15};
        for (int i : zero)
            if (P[i] != 0.0f) return fail(ctx, CR_EUNSUPPORTED, "cr_cone_trace: P is not a symmetric perspective projection");
        if (P[11] != -1.0f || !(P[0] > 0.0f) || !(P[5] > 0.0f))
            return fail(ctx, CR_EUNSUPPORTED, "cr_cone_trace: P is not a symmetric perspective projection");
        for (int i = 0; i < 16; ++i) cv.V[i] = cam->V[i];
        cv.kx = 1.0f / P[0];
        cv.ky = 1.0f / P[5];
        cv.znear = P[14] / (P[10] - 1.0f);
        cv.zfar = P[14] / (P[10] + 1.0f);
        if (!(cv.znear > 0.0f) || !(cv.zfar > cv.znear) || !std::isfinite(cv.zfar))
            return fail(ctx, CR_EINVAL, "cr_cone_trace: P gives no 0 < znear < zfar");
        float r0 = cv.V[0] * cv.V[0] + cv.V[4] * cv.V[4] + cv.V[8] * cv.V[8];
        float r2 = cv.V[2] * cv.V[2] + cv.V[6] * cv.V[6] + cv.V[10] * cv.V[10];
        if (!(r0 > 0.0f) || !(r2 > 0.0f)) return fail(ctx, CR_EINVAL, "cr_cone_trace: degenerate view matrix");
    } else {
        if (!finite3(cam->target) || !finite3(cam->up)) return fail(ctx, CR_EINVAL, "cr_cone_trace: non-finite camera");
        float tanHalfFovy = std::tan(cam->fovy / 2.0f);
        if (!(tanHalfFovy > 0.0f) || !std::isfinite(tanHalfFovy))
            return fail(ctx, CR_EINVAL, "cr_cone_trace: fovy needs tan(fovy / 2) > 0");
        if (!(cam->aspect >= 0.0f) || !std::isfinite(cam->aspect))
            return fail(ctx, CR_EINVAL, "cr_cone_trace: aspect must be >= 0 (0: width / height)");
        if (!(cam->znear > 0.0f) || !(cam->zfar > cam->znear)) return fail(ctx, CR_EINVAL, "cr_cone_trace: need 0 < znear < zfar");
        F3 e = f3(cam->eye[0], cam->eye[1], cam->eye[2]), t = f3(cam->target[0], cam->target[1], cam->target[2]);
        F3 d = t - e, upv = f3(cam->up[0], cam->up[1], cam->up[2]);
        F3 sd = cross3(d, upv);
        if (!(dot3(d, d) > 0.0f) || !(dot3(sd, sd) > 0.0f))
            return fail(ctx, CR_EINVAL, "cr_cone_trace: eye, target and up give no view basis");
        look_at(e, t, upv, cv.V);
        float aspect = cam->aspect > 0.0f ? cam->aspect : (float)W / (float)H;
        cv.kx = tanHalfFovy * aspect;
        cv.ky = tanHalfFovy;
        cv.znear = cam->znear;
        cv.zfar = cam->zfar;
    }
    basis_from_view(cv.V, cv.S, cv.U, cv.B);
    for (int a = 0; a < 3; ++a) cv.eye[a] = cam->eye[a];
    return CR_OK;
}

// validation, camera view, cone schedule and buffer allocation (tile_w x
// 64 / tile_w camera tiles); the volume's voxelized state is not checked here
static int cone_prepare(cr_volume v, const cr_light* light, const cr_camera* cam, const cr_cone_params* p,
                        const LearnedOpts* lo, void* d_image, int W, int H, size_t pitch, int pixfmt, int tile_w,
                        ConePlan& plan) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_cone_trace: vol is NULL");
    cr_context ctx = v->ctx;
    if (!light || !cam || !p || !d_image) return fail(ctx, CR_EINVAL, "cr_cone_trace: NULL argument");
    if (v->sun_on && p->composite != CR_COMPOSITE_GL_BLEND)
        return fail(ctx, CR_EUNSUPPORTED, "cr_cone_trace: the sun sprite needs the GL-blend composite");
    if (W <= 0 || H <= 0 || W > 16384 || H > 16384) return fail(ctx, CR_EINVAL, "cr_cone_trace: image size must be in [1, 16384]");
    if (pixfmt != CR_PIXEL_RGBA32F && pixfmt != CR_PIXEL_RGBA8) return fail(ctx, CR_EINVAL, "cr_cone_trace: bad pixel format");
    size_t px_bytes = pixfmt == CR_PIXEL_RGBA32F ? 16 : 4;
    if (pitch < (size_t)W * px_bytes) return fail(ctx, CR_EINVAL, "cr_cone_trace: pitch smaller than a row");
    if (pitch % px_bytes != 0 || ((uintptr_t)d_image % px_bytes) != 0)
        return fail(ctx, CR_EINVAL, "cr_cone_trace: image pointer or pitch misaligned");
    if (tile_w != 4 && tile_w != 8 && tile_w != 16 && tile_w != 32)
        return fail(ctx, CR_EINVAL, "cr_cone_trace: tile width must be 4, 8, 16 or 32");
    int rc = cone_params_check(ctx, v, p, "cr_cone_trace");
    if (rc) return rc;
    if (lo) {
        if (!lo->mlp) return fail(ctx, CR_EINVAL, "cr_cone_trace_learned: mlp is NULL");
        if (p->cone_mode == CR_CONE_WEIGHTED || p->composite != CR_COMPOSITE_FTB)
            return fail(ctx, CR_EUNSUPPORTED, "cr_cone_trace_learned: needs a front-to-back cone and composite");
        if (lo->k < 1 || lo->k > kMaxSteps) return fail(ctx, CR_EINVAL, "cr_cone_trace_learned: k must be in [1, 256]");
        if (std::isnan(lo->err_bound)) return fail(ctx, CR_EINVAL, "cr_cone_trace_learned: error bound is NaN");
    }
    if (light->directional && !finite3(light->direction)) return fail(ctx, CR_EINVAL, "cr_cone_trace: non-finite light direction");
    if (!light->directional && !finite3(light->position)) return fail(ctx, CR_EINVAL, "cr_cone_trace: non-finite light position");
    CameraView& cv = plan.cv;
    cv = CameraView{};
    if ((rc = camera_view(ctx, cam, W, H, cv))) return rc;
    rc = set_device(ctx);
    if (rc) return rc;
    cv.W = W;
    cv.H = H;
    cv.tile_w = tile_w;
    cv.tile_h = 64 / tile_w;
    cv.tile_shift = tile_w == 4 ? 2 : tile_w == 8 ? 3 : tile_w == 16 ? 4 : 5;
    cv.tiles_x = (W + cv.tile_w - 1) / cv.tile_w;
    cv.tiles_y = (H + cv.tile_h - 1) / cv.tile_h;

    // cone schedule (res/conetrace_frag.glsl:70-76), identical for every ray
    ConeDesc& cd = plan.cd;
    cd = ConeDesc{};
    fill_cone_desc(v, light, p, cd);
    cd.pixfmt = pixfmt;
    cd.pitch = pitch;
    if (lo) {
        cd.learned_k = lo->k;
        cd.err_bound = lo->err_bound;
    }

    // buffers
    int S = v->S;
    plan.S = S;
    size_t scap = (size_t)v->sph_cap;
    if (S > 0) {
        size_t c0 = scap, c1 = scap, c2 = scap, c3 = scap;
        if ((rc = ensure(ctx, &v->d_sph_world, (size_t)S, &c0))) return rc;
        if ((rc = ensure(ctx, &v->d_sph_view, (size_t)S, &c1))) return rc;
        if ((rc = ensure(ctx, &v->d_sph_tiles, (size_t)S, &c2))) return rc;
        if ((rc = ensure(ctx, &v->d_order, (size_t)S, &c3))) return rc;
        v->sph_cap = (int)std::min(std::min(c0, c1), std::min(c2, c3));
        if ((rc = ensure(ctx, &v->d_sort, sort_scratch_words(S), &v->sort_cap))) return rc;
    }
    plan.words = (S + 63) / 64;
    plan.ntiles = (size_t)cv.tiles_x * cv.tiles_y;
    plan.nmask = plan.ntiles * (size_t)std::max(plan.words, 1);
    if ((rc = ensure(ctx, &v->d_masks, plan.nmask, &v->masks_cap))) return rc;
    if ((rc = ensure(ctx, &v->d_tiles, tile_scratch_ints((int)plan.ntiles), &v->tiles_cap))) return rc;
    return CR_OK;
}

// camera setup + draw order (kernels only)
// (CR_ORDER_UPLOAD: the billboards are drawn in upload order, nothing sorted)
static void cone_enqueue_sort(cr_volume v, const ConePlan& plan, StageEvents* ev, hipStream_t s) {
    launch_cam_setup(v->g, plan.cv, v->d_boards, plan.S, v->d_sph_world, v->d_sph_view, v->d_sph_tiles, v->d_sort, s);
    if (ev) ev->mark(1);
    if (v->order_mode == CR_ORDER_UPLOAD) launch_iota(v->d_order, plan.S, s);
    else launch_sort(v->d_sort, plan.S, v->d_order, v->d_tie, s);
    if (ev) ev->mark(2);
}

// exchange the volume's camera buffers with the spare pair
static void swap_cam(cr_volume v) {
    CamBufs& o = v->cam_spare;
    std::swap(v->d_sph_world, o.sph_world);
    std::swap(v->d_sph_view, o.sph_view);
    std::swap(v->d_sph_tiles, o.sph_tiles);
    std::swap(v->d_sort, o.sort);
    std::swap(v->d_order, o.order);
    std::swap(v->d_tie, o.tie);
    std::swap(v->d_tiles, o.tiles);
    std::swap(v->d_masks, o.masks);
    std::swap(v->sph_cap, o.sph_cap);
    std::swap(v->sort_cap, o.sort_cap);
    std::swap(v->masks_cap, o.masks_cap);
    std::swap(v->tiles_cap, o.tiles_cap);
    v->cset ^= 1;
}

static int cone_trace_impl(cr_volume v, const cr_light* light, const cr_camera* cam, const cr_cone_params* p,
                           const LearnedOpts* lo, void* d_image, int W, int H, size_t pitch, int pixfmt, void* stream) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_cone_trace: vol is NULL");
    cr_context ctx = v->ctx;
    // Camera setup, sort and binning depend on the billboards and the camera
    // only: they run on the volume's camera-side stream into the buffer pair
    // the previous trace did not use, from the end of the trace before it
    // on, so they overlap the previous cone trace and the voxelization
    // enqueued before this call; the cone kernel waits for them.  Counter
    // runs stay serial.
    const bool overlap = ctx->profiling != 1;
    bool swapped = false;
    if (overlap && v->voxelized) {
        if (!v->cam_spare.tie && hipMalloc((void**)&v->cam_spare.tie, 64) != hipSuccess) {
            (void)hipGetLastError();  // no second pair: the camera pass waits for the previous trace
            v->cam_spare.tie = nullptr;
        }
        if (v->cam_spare.tie) {
            swap_cam(v);
            swapped = true;
        }
    }
    ConePlan plan;
    int rc = cone_prepare(v, light, cam, p, lo, d_image, W, H, pitch, pixfmt, kTile, plan);
    if (rc) {
        if (swapped) swap_cam(v);
        return rc;
    }
    if (!v->voxelized) return fail(ctx, CR_ESTATE, "cr_cone_trace: volume not voxelized");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* counters = ctx->profiling == 1 ? v->d_counters : nullptr;
    hipStream_t cs = s;
    if (overlap) {
        if (!v->side) {
            v->side = ctx->lside;
            for (hipEvent_t* e : {&v->ev_cam, &v->ev_csync, &v->ev_cfree[0], &v->ev_cfree[1]})
                CR_HIP(ctx, hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
        cs = v->side;
        if (v->cam_sync_needed) {  // after all prior work on s
            CR_HIP(ctx, hipEventRecord(v->ev_csync, s));
            CR_HIP(ctx, hipStreamWaitEvent(cs, v->ev_csync, 0));
            v->cam_sync_needed = false;
        }
        if (v->cfree_valid[v->cset]) CR_HIP(ctx, hipStreamWaitEvent(cs, v->ev_cfree[v->cset], 0));
    }
    StageEvents ev;
    if ((rc = ev.begin(ctx, cs, 1))) return rc;
    ev.mark(0);
    if (counters) launch_fill_u64((uint64_t*)counters, 9, 0, s);
    cone_enqueue_sort(v, plan, &ev, cs);
    TileQueue tq = launch_bin(v->d_sph_tiles, v->d_order, plan.S, plan.cv.tiles_x, plan.cv.tiles_y, plan.words,
                              v->d_masks, v->d_tiles, cs);
    ev.mark(3);
    if (overlap) {
        CR_HIP(ctx, hipEventRecord(v->ev_cam, cs));
        CR_HIP(ctx, hipStreamWaitEvent(s, v->ev_cam, 0));
    }
    ev.mark_on(4, s);
    ConeLaunch cl{};
    cl.g = v->g;
    cl.vd = v->vd;
    cl.cv = plan.cv;
    cl.cd = plan.cd;
    cl.cd.count = counters ? 1 : 0;  // the march counts are read only from counter runs
    cl.levels = v->d_levels;
    cl.sph_world = v->d_sph_world;
    cl.sph_view = v->d_sph_view;
    cl.order = v->d_order;
    cl.masks = v->d_masks;
    cl.words = plan.words;
    cl.image = d_image;
    cl.counters = counters;
    cl.tile_order = tq.order;
    cl.learned = lo ? 1 : 0;
    cl.table_addressing = (ctx->flags & CR_CONTEXT_DIRECT_ADDRESSING) ? 0 : 1;
    if (lo) cl.mlp = mlp_dev(lo->mlp);
    if (v->sun_on) {
        // SunShader::render before the clouds: centre Sun::position, clipped
        // like any primitive by the camera's depth range
        const float* Vm = cl.cv.V;
        const float px = light->position[0], py = light->position[1], pz = light->position[2];
        const float vx = ((Vm[0] * px + Vm[4] * py) + Vm[8] * pz) + Vm[12];
        const float vy = ((Vm[1] * px + Vm[5] * py) + Vm[9] * pz) + Vm[13];
        const float vz = ((Vm[2] * px + Vm[6] * py) + Vm[10] * pz) + Vm[14];
        const float dz = -vz;
        if (dz >= plan.cv.znear && dz <= plan.cv.zfar) {
            ConeDesc& cd = cl.cd;
            cd.sun_on = 1;
            cd.sun_v[0] = vx; cd.sun_v[1] = vy; cd.sun_v[2] = vz;
            for (int a = 0; a < 3; ++a) { cd.sun_in[a] = v->sun.inner_color[a]; cd.sun_out[a] = v->sun.outer_color[a]; }
            cd.sun_r0 = v->sun.inner_radius;
            cd.sun_r1 = v->sun.outer_radius;
        }
    }
    if (v->nm_on) {
        ConeDesc& cd = cl.cd;
        cd.nm_octaves = v->nm.octaves;
        cd.nm_min_steps = v->nm.min_steps;
        cd.nm_max_steps = v->nm.max_steps;
        cd.nm_step = v->nm.step_size;
        cd.nm_opacity = v->nm.opacity;
        cd.nm_freq_step = v->nm.freq_step;
        cd.nm_pers_step = v->nm.pers_step;
        cd.nm_adjust = v->nm.adjust_size;
        cd.nm_min_color = v->nm.min_color;
        cd.nm_color_scale = v->nm.color_scale;
        for (int a = 0; a < 3; ++a) cd.nm_off[a] = v->nm.offsets[a];
        // viewRay = normalize(vec3(V[0][2], V[1][2], V[2][2])) (res/conetrace_frag.glsl:138)
        const F3 ray = norm3(f3(cl.cv.V[2], cl.cv.V[6], cl.cv.V[10]));
        cd.nm_ray[0] = ray.x; cd.nm_ray[1] = ray.y; cd.nm_ray[2] = ray.z;
        cl.nmap = v->d_nmap;
    }
    launch_cone(cl, s);
    ev.mark_on(5, s);
    ev.end();
    if (overlap) {
        CR_HIP(ctx, hipEventRecord(v->ev_cfree[v->cset], s));
        v->cfree_valid[v->cset] = true;
    } else {
        v->cam_sync_needed = true;
    }
    CR_HIP(ctx, hipGetLastError());
    v->sorted_count = plan.S;
    if (ctx->profiling == 1) {
        CR_HIP(ctx, hipStreamSynchronize(s));
        unsigned long long c[9];
        CR_HIP(ctx, hipMemcpy(c, v->d_counters, sizeof(c), hipMemcpyDeviceToHost));
        ctx->stats.mlp_rows = lo ? (int64_t)c[6] : 0;
        ctx->stats.mlp_fallbacks = lo ? (int64_t)c[7] : 0;
        ctx->stats.mlp_evals = lo ? (int64_t)c[8] : 0;
        ctx->stats.fragments = (int64_t)c[3];
        ctx->stats.samples = (int64_t)c[4];
        ctx->stats.two_level = (int64_t)c[5];
        ctx->stats.ms_setup = elapsed(ev.e[0], ev.e[1]);
        ctx->stats.ms_sort = elapsed(ev.e[1], ev.e[2]);
        ctx->stats.ms_bin = elapsed(ev.e[2], ev.e[3]);
        ctx->stats.ms_cone = elapsed(ev.e[4], ev.e[5]);
        ctx->stats.n_cone = 1;
    }
    return CR_OK;
}

int cr_cone_trace(cr_volume v, const cr_light* light, const cr_camera* cam, const cr_cone_params* p, void* d_image,
                  int W, int H, size_t pitch, int pixfmt, void* stream) {
    return cone_trace_impl(v, light, cam, p, nullptr, d_image, W, H, pitch, pixfmt, stream);
}

int cr_cone_trace_learned(cr_volume v, const cr_light* light, const cr_camera* cam, const cr_cone_params* p,
                          cr_mlp mlp, int k, float err_bound, void* d_image, int W, int H, size_t pitch, int pixfmt,
                          void* stream) {
    if (v && v->nm_on)
        return fail(v->ctx, CR_EUNSUPPORTED, "cr_cone_trace_learned: the noise march is on for this volume");
    LearnedOpts lo{mlp, k, err_bound};
    return cone_trace_impl(v, light, cam, p, &lo, d_image, W, H, pitch, pixfmt, stream);
}

int cr_volume_segments(cr_volume v, const cr_light* light, const cr_cone_params* p, int k, const float* d_p0, int n,
                       float* d_features, float* d_targets, int32_t* d_valid, void* stream) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_segments: vol is NULL");
    cr_context ctx = v->ctx;
    if (!light || !p || (n > 0 && (!d_p0 || !d_features || !d_targets || !d_valid)))
        return fail(ctx, CR_EINVAL, "cr_volume_segments: NULL argument");
    if (n < 0) return fail(ctx, CR_EINVAL, "cr_volume_segments: n must be >= 0");
    if (v->nm_on) return fail(ctx, CR_EUNSUPPORTED, "cr_volume_segments: the noise march is on for this volume");
    int rc = cone_params_check(ctx, v, p, "cr_volume_segments");
    if (rc) return rc;
    if (p->cone_mode == CR_CONE_WEIGHTED) return fail(ctx, CR_EUNSUPPORTED, "cr_volume_segments: needs a front-to-back cone");
    if (!light->directional || !finite3(light->direction))
        return fail(ctx, CR_EUNSUPPORTED, "cr_volume_segments: needs a finite directional light");
    if (k < 1 || k > kMaxSteps) return fail(ctx, CR_EINVAL, "cr_volume_segments: k must be in [1, 256]");
    if (!v->voxelized) return fail(ctx, CR_ESTATE, "cr_volume_segments: volume not voxelized");
    if ((rc = set_device(ctx))) return rc;
    ConeDesc cd{};
    fill_cone_desc(v, light, p, cd);
    cd.learned_k = k;
    launch_segments(v->g, v->vd, cd, v->d_levels, d_p0, n, d_features, d_targets, (int*)d_valid, (hipStream_t)stream);
    CR_HIP(ctx, hipGetLastError());
    return CR_OK;
}

int cr_volume_occupancy(cr_volume v, uint64_t* count, uint64_t* hash, void* stream) {
    if (!v || !count || !hash) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_occupancy: NULL argument");
    cr_context ctx = v->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_fill_u64((uint64_t*)v->d_counters, 2, 0, s);
    launch_occupancy(v->g, v->d_shell, v->mode == CR_MODE_RGBA ? v->d_fill : nullptr, v->d_counters, s);
    CR_HIP(ctx, hipGetLastError());
    unsigned long long c[2];
    CR_HIP(ctx, hipMemcpyAsync(c, v->d_counters, sizeof(c), hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    *count = c[0];
    *hash = c[1];
    return CR_OK;
}

// VoxelShader's activeVoxels (src/Shaders/VoxelShader.cpp:110-115): level-0
// texels with r != 0, which are exactly the lit-shell voxels in every mode
// (chord-fill texels are (0, 0, 0, 1); noise scales a only)
int cr_volume_active_voxels(cr_volume v, uint64_t* count, void* stream) {
    if (!v || !count) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_active_voxels: NULL argument");
    cr_context ctx = v->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_fill_u64((uint64_t*)v->d_counters, 2, 0, s);
    launch_occupancy(v->g, v->d_shell, nullptr, v->d_counters, s);
    CR_HIP(ctx, hipGetLastError());
    unsigned long long c[2];
    CR_HIP(ctx, hipMemcpyAsync(c, v->d_counters, sizeof(c), hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    *count = c[0];
    return CR_OK;
}

static int readback_scratch(cr_volume v, size_t bytes) {
    size_t cap = v->scratch_cap;
    int rc = ensure(v->ctx, &v->d_scratch, bytes, &cap);
    v->scratch_cap = cap;
    return rc;
}

int cr_volume_read_level(cr_volume v, int level, void* host_dst, size_t bytes, void* stream) {
    if (!v || !host_dst) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_read_level: NULL argument");
    cr_context ctx = v->ctx;
    if (level < 0 || level >= v->L) return fail(ctx, CR_EINVAL, "cr_volume_read_level: level out of range");
    uint64_t n = v->vd.lv[level].n;
    size_t need = (size_t)(n * n * n) * texel_bytes(v->format);
    if (bytes < need) return fail(ctx, CR_EINVAL, "cr_volume_read_level: destination too small");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = readback_scratch(v, need))) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_debrick(v->vd, level, v->d_levels, v->d_scratch, s);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipMemcpyAsync(host_dst, v->d_scratch, need, hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

int cr_volume_read_occupancy(cr_volume v, uint8_t* host_dst, size_t bytes, void* stream) {
    if (!v || !host_dst) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_read_occupancy: NULL argument");
    cr_context ctx = v->ctx;
    size_t need = (size_t)v->N * v->N * v->N;
    if (bytes < need) return fail(ctx, CR_EINVAL, "cr_volume_read_occupancy: destination too small");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = readback_scratch(v, need))) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_read_state(v->g, v->d_shell, v->mode == CR_MODE_RGBA ? v->d_fill : nullptr, v->d_scratch, s);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipMemcpyAsync(host_dst, v->d_scratch, need, hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

int cr_volume_read_position_map(cr_volume v, float* host_dst, size_t bytes, void* stream) {
    if (!v || !host_dst) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_read_position_map: NULL argument");
    cr_context ctx = v->ctx;
    if (!v->voxelized) return fail(ctx, CR_ESTATE, "cr_volume_read_position_map: volume not voxelized");
    if (v->gen != v->vox_gen)
        return fail(ctx, CR_ESTATE, "cr_volume_read_position_map: billboards changed since the last voxelize");
    size_t npx = (size_t)v->lw * v->lh;
    size_t need = npx * sizeof(float4);
    if (bytes < need) return fail(ctx, CR_EINVAL, "cr_volume_read_position_map: destination too small");
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = ensure(ctx, &v->d_posmap, npx, &v->posmap_cap))) return rc;
    launch_light_splat(v->g, v->last_lv, v->d_boards, v->d_keys, v->d_posmap, nullptr, nullptr, s);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipMemcpyAsync(host_dst, v->d_posmap, need, hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

int cr_volume_set_draw_order(cr_volume v, int order_mode) {
    if (!v) return fail(nullptr, CR_EINVAL, "cr_volume_set_draw_order: vol is NULL");
    if (order_mode != CR_ORDER_SORTED && order_mode != CR_ORDER_UPLOAD)
        return fail(v->ctx, CR_EINVAL, "cr_volume_set_draw_order: mode must be CR_ORDER_SORTED or CR_ORDER_UPLOAD");
    if (order_mode != v->order_mode) v->gen += 1;  // batch graphs bake the order in
    v->order_mode = order_mode;
    return CR_OK;
}

// CloudVolume::sortBoards(point) on the device, in buffers of its own (a
// queued cone trace keeps reading its camera pair's draw order)
int cr_volume_sort_boards(cr_volume v, const float point[3], int32_t* host_order, int count, void* stream) {
    if (!v || !point || (!host_order && count > 0))
        return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_sort_boards: NULL argument");
    cr_context ctx = v->ctx;
    if (!finite3(point)) return fail(ctx, CR_EINVAL, "cr_volume_sort_boards: non-finite point");
    if (count < v->S) return fail(ctx, CR_EINVAL, "cr_volume_sort_boards: count smaller than the billboard count");
    if (v->S == 0) return CR_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure(ctx, &v->d_xsort, sort_scratch_words(v->S), &v->xsort_cap))) return rc;
    if ((rc = ensure(ctx, &v->d_xorder, (size_t)v->S, &v->xorder_cap))) return rc;
    if ((rc = ensure(ctx, &v->d_xtie, (size_t)16, &v->xtie_cap))) return rc;  // 64 bytes
    hipStream_t s = (hipStream_t)stream;
    launch_sort_keys(v->g, v->d_boards, v->S, point, v->d_xsort, s);
    launch_sort(v->d_xsort, v->S, v->d_xorder, v->d_xtie, s);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipMemcpyAsync(host_order, v->d_xorder, (size_t)v->S * 4, hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

int cr_volume_read_draw_order(cr_volume v, int32_t* host_dst, int count, void* stream) {
    if (!v || !host_dst) return fail(v ? v->ctx : nullptr, CR_EINVAL, "cr_volume_read_draw_order: NULL argument");
    cr_context ctx = v->ctx;
    if (v->sorted_count < 0) return fail(ctx, CR_ESTATE, "cr_volume_read_draw_order: no cone trace since the last upload");
    if (count < v->sorted_count) return fail(ctx, CR_EINVAL, "cr_volume_read_draw_order: destination too small");
    if (v->sorted_count == 0) return CR_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    CR_HIP(ctx, hipMemcpyAsync(host_dst, v->d_order, (size_t)v->sorted_count * 4, hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// far-field MLP
// ---------------------------------------------------------------------------
extern "C" {

int cr_mlp_create(cr_context ctx, cr_mlp* out) {
    if (!ctx || !out) return fail(ctx, CR_EINVAL, "cr_mlp_create: NULL argument");
    *out = nullptr;
    int rc = set_device(ctx);
    if (rc) return rc;
    cr_mlp m = new (std::nothrow) cr_mlp_s();
    if (!m) return fail(ctx, CR_ENOMEM, "cr_mlp_create: host allocation failed");
    m->ctx = ctx;
    hipError_t e = hipMalloc((void**)&m->d_master, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_m1, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_m2, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_grad, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_fwd, kMlpFwdBlocks * 512 * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_bwd, kMlpBwdBlocks * 512 * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_bias, kMlpBiasPad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_loss, sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->d_master, 0, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->d_m1, 0, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->d_m2, 0, kMlpParamsPad * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->d_loss, 0, sizeof(float));
    if (e != hipSuccess) {
        cr_mlp_destroy(m);
        return fail(ctx, e == hipErrorOutOfMemory ? CR_ENOMEM : CR_EDEVICE, "cr_mlp_create: %s", hipGetErrorString(e));
    }
    launch_mlp_pack(m->d_master, m->d_fwd, m->d_bwd, m->d_bias, nullptr);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipDeviceSynchronize());
    *out = m;
    return CR_OK;
}

int cr_mlp_destroy(cr_mlp m) {
    if (!m) return CR_OK;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    void* ptrs[] = {m->d_master, m->d_m1, m->d_m2, m->d_grad, m->d_fwd, m->d_bwd,
                    m->d_bias, m->d_loss, m->d_partial, m->d_loss_part};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete m;
    return CR_OK;
}

int cr_mlp_set_params(cr_mlp m, const float* host, int count) {
    if (!m || !host) return fail(m ? m->ctx : nullptr, CR_EINVAL, "cr_mlp_set_params: NULL argument");
    cr_context ctx = m->ctx;
    if (count != kMlpParams) return fail(ctx, CR_EINVAL, "cr_mlp_set_params: count must be CR_MLP_PARAMS");
    for (int i = 0; i < count; ++i)
        if (!std::isfinite(host[i])) return fail(ctx, CR_EINVAL, "cr_mlp_set_params: non-finite parameter");
    int rc = set_device(ctx);
    if (rc) return rc;
    CR_HIP(ctx, hipMemcpy(m->d_master, host, (size_t)count * sizeof(float), hipMemcpyHostToDevice));
    CR_HIP(ctx, hipMemset(m->d_m1, 0, kMlpParamsPad * sizeof(float)));
    CR_HIP(ctx, hipMemset(m->d_m2, 0, kMlpParamsPad * sizeof(float)));
    m->step = 0;
    launch_mlp_pack(m->d_master, m->d_fwd, m->d_bwd, m->d_bias, nullptr);
    CR_HIP(ctx, hipGetLastError());
    CR_HIP(ctx, hipDeviceSynchronize());
    return CR_OK;
}

int cr_mlp_get_params(cr_mlp m, float* host, int count) {
    if (!m || !host) return fail(m ? m->ctx : nullptr, CR_EINVAL, "cr_mlp_get_params: NULL argument");
    cr_context ctx = m->ctx;
    if (count != kMlpParams) return fail(ctx, CR_EINVAL, "cr_mlp_get_params: count must be CR_MLP_PARAMS");
    int rc = set_device(ctx);
    if (rc) return rc;
    CR_HIP(ctx, hipDeviceSynchronize());
    CR_HIP(ctx, hipMemcpy(host, m->d_master, (size_t)count * sizeof(float), hipMemcpyDeviceToHost));
    return CR_OK;
}

int cr_mlp_forward(cr_mlp m, const float* d_x, int n, float* d_y, void* stream) {
    if (!m) return fail(nullptr, CR_EINVAL, "cr_mlp_forward: mlp is NULL");
    cr_context ctx = m->ctx;
    if (n < 0 || (n > 0 && (!d_x || !d_y))) return fail(ctx, CR_EINVAL, "cr_mlp_forward: bad arguments");
    int rc = set_device(ctx);
    if (rc) return rc;
    launch_mlp_forward(mlp_dev(m), d_x, n, d_y, (hipStream_t)stream);
    CR_HIP(ctx, hipGetLastError());
    return CR_OK;
}

// forward + backward over n rows into the gradient partials of ceil(n / 64)
// workgroups (each writes all of its own), reduced in a fixed order
// (deterministic), optionally followed by an Adam step; kernels only
static int mlp_launch_step(cr_mlp m, const float* d_x, const float* d_t, int n, float lr, int apply, float* partial,
                           float* loss_part, hipStream_t s) {
    cr_context ctx = m->ctx;
    const int nwg = mlp_train_groups(n);
    launch_mlp_train(mlp_dev(m), d_x, d_t, n, partial, loss_part, s);
    AdamStep st{};
    st.apply = apply;
    st.lr = lr;
    st.beta1 = 0.9f;
    st.beta2 = 0.999f;
    st.eps = 1e-8f;
    if (apply) {
        m->step += 1;
        st.bc1 = (float)(1.0 - std::pow(0.9, (double)m->step));
        st.bc2 = (float)(1.0 - std::pow(0.999, (double)m->step));
    } else {
        st.bc1 = st.bc2 = 1.0f;
    }
    launch_mlp_adam(partial, nwg, loss_part, m->d_master, m->d_m1, m->d_m2, m->d_grad, m->d_loss, st, s);
    if (apply) launch_mlp_pack(m->d_master, m->d_fwd, m->d_bwd, m->d_bias, s);
    CR_HIP(ctx, hipGetLastError());
    return CR_OK;
}

// mlp_launch_step with partials in buffers of the cr_mlp, grown on demand
static int mlp_step(cr_mlp m, const float* d_x, const float* d_t, int n, float lr, int apply, hipStream_t s) {
    cr_context ctx = m->ctx;
    const int nwg = mlp_train_groups(n);
    int rc = ensure(ctx, &m->d_partial, (size_t)nwg * kMlpParamsPad, &m->partial_cap);
    if (rc) return rc;
    if ((rc = ensure(ctx, &m->d_loss_part, (size_t)nwg, &m->loss_cap))) return rc;
    return mlp_launch_step(m, d_x, d_t, n, lr, apply, m->d_partial, m->d_loss_part, s);
}

int cr_mlp_gradient(cr_mlp m, const float* d_x, const float* d_t, int n, float* host_grad, float* host_loss,
                    void* stream) {
    if (!m) return fail(nullptr, CR_EINVAL, "cr_mlp_gradient: mlp is NULL");
    cr_context ctx = m->ctx;
    if (n < 1 || !d_x || !d_t || !host_grad || !host_loss) return fail(ctx, CR_EINVAL, "cr_mlp_gradient: bad arguments");
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = mlp_step(m, d_x, d_t, n, 0.0f, 0, s))) return rc;
    CR_HIP(ctx, hipMemcpyAsync(host_grad, m->d_grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipMemcpyAsync(host_loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

int cr_mlp_train_step(cr_mlp m, const float* d_x, const float* d_t, int n, float lr, void* stream) {
    if (!m) return fail(nullptr, CR_EINVAL, "cr_mlp_train_step: mlp is NULL");
    cr_context ctx = m->ctx;
    if (n < 1 || !d_x || !d_t) return fail(ctx, CR_EINVAL, "cr_mlp_train_step: bad arguments");
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(ctx, CR_EINVAL, "cr_mlp_train_step: bad learning rate");
    int rc = set_device(ctx);
    if (rc) return rc;
    return mlp_step(m, d_x, d_t, n, lr, 1, (hipStream_t)stream);
}

// cr_mlp_train_step with the partials in this device's `scratch`: no
// allocation (the pointer was looked up when the context was created)
int cr_mlp_train_step_static(cr_mlp m, const float* d_x, const float* d_t, int n, float lr, void* stream) {
    if (!m) return fail(nullptr, CR_EINVAL, "cr_mlp_train_step_static: mlp is NULL");
    cr_context ctx = m->ctx;
    if (!d_x || !d_t) return fail(ctx, CR_EINVAL, "cr_mlp_train_step_static: NULL array");
    if (n < 1 || n > CR_MLP_STATIC_ROWS) {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "n = %d outside [1, CR_MLP_STATIC_ROWS = %d]", n, CR_MLP_STATIC_ROWS);
        return fail(ctx, CR_EINVAL, "cr_mlp_train_step_static: %s", msg);
    }
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(ctx, CR_EINVAL, "cr_mlp_train_step_static: bad learning rate");
    if (!ctx->static_scratch) return fail(ctx, CR_ESTATE, "cr_mlp_train_step_static: no scratch array on this device");
    int rc = set_device(ctx);
    if (rc) return rc;
    float* partial = ctx->static_scratch;
    return mlp_launch_step(m, d_x, d_t, n, lr, 1, partial, partial + kStaticPartials, (hipStream_t)stream);
}

int cr_mlp_last_loss(cr_mlp m, float* loss, void* stream) {
    if (!m || !loss) return fail(m ? m->ctx : nullptr, CR_EINVAL, "cr_mlp_last_loss: NULL argument");
    cr_context ctx = m->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    CR_HIP(ctx, hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
    CR_HIP(ctx, hipStreamSynchronize(s));
    return CR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched scenes: one hipGraph of every scene's voxelize / sort / bin and a
// single cone-trace launch over all scenes' tiles
// ---------------------------------------------------------------------------
struct cr_batch_s {
    cr_context ctx = nullptr;
    std::vector<cr_batch_item> items;
    int format = 0, cone_mode = 0, composite = 0;
    cr_batch_config cfg{8, 1, 0};
    BatchScene* d_scenes = nullptr;
    size_t scenes_cap = 0;
    int* d_tile0 = nullptr;
    size_t tile0_cap = 0;
    int* d_counts = nullptr;
    size_t counts_cap = 0;
    int* d_order = nullptr;
    size_t order_cap = 0;
    int* d_oscratch = nullptr;
    size_t oscratch_cap = 0;
    unsigned long long* d_counters = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_graph = false;
    cr_batch_config graph_cfg{0, 0, 0};
    std::vector<uint64_t> graph_gen;
    std::vector<const void*> graph_ptrs;  // per scene: the camera and work-list buffers the graph was captured with
    int total_tiles = 0;
    // batched stage launches (scenes batch_vox_eligible takes): per-scene
    // descriptors, the stages' first-workgroup and scene tables, fill segments
    BatchVox* d_vox = nullptr;
    size_t vox_cap = 0;
    int* d_stage = nullptr;
    size_t stage_cap = 0;
    FillSeg* d_segs = nullptr;
    size_t segs_cap = 0;
    struct StageLaunch {
        int stage, variant, first_off, n, blocks;
    };
    std::vector<StageLaunch> launches;
    std::vector<char> batched;  // per scene: voxelization and camera pass in the batched stages
    int fill_off = 0, nseg = 0, fill_blocks = 0;
};

namespace {

bool same_cfg(const cr_batch_config& a, const cr_batch_config& b) {
    return a.tile_w == b.tile_w && a.waves_per_group == b.waves_per_group && a.groups == b.groups;
}

void drop_graph(cr_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->have_graph = false;
}

// host preparation (allocations, descriptor upload), then the kernels
int batch_prepare(cr_batch b, std::vector<VoxPlan>& vp, std::vector<ConePlan>& cp) {
    cr_context ctx = b->ctx;
    const size_t n = b->items.size();
    vp.resize(n);
    cp.resize(n);
    std::vector<BatchScene> hs(n);
    std::vector<int> tile0(n);
    int total = 0;
    for (size_t i = 0; i < n; ++i) {
        const cr_batch_item& it = b->items[i];
        cr_volume v = it.volume;
        int rc = voxelize_prepare(v, &it.light, it.light_w, it.light_h, vp[i]);
        if (rc) return rc;
        if ((rc = light_scratch(v, vp[i]))) return rc;
        rc = cone_prepare(v, &it.light, &it.camera, &it.params, nullptr, it.d_image, it.width, it.height,
                          it.pitch_bytes, it.pixel_format, b->cfg.tile_w, cp[i]);
        if (rc) return rc;
        if (cp[i].ntiles > (size_t)(1 << 30) - (size_t)total)
            return fail(ctx, CR_EUNSUPPORTED, "cr_batch: too many tiles");
        BatchScene& B = hs[i];
        B.g = v->g;
        B.vd = v->vd;
        B.cv = cp[i].cv;
        B.cd = cp[i].cd;
        B.levels = v->d_lset[0];  // batches build and read set 0
        B.sph_world = v->d_sph_world;
        B.sph_view = v->d_sph_view;
        B.order = v->d_order;
        B.masks = v->d_masks;
        B.image = it.d_image;
        B.words = cp[i].words;
        B.tile0 = total;
        tile0[i] = total;
        total += (int)cp[i].ntiles;
    }
    b->total_tiles = total;
    int rc;
    // batched stages: descriptors of the scenes they take, then per stage
    // (and launch variant) the members' first workgroups and scene indices
    std::vector<BatchVox> hv(n);
    b->batched.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        cr_volume v = b->items[i].volume;
        if (!batch_vox_eligible(v->g, v->vd, v->S)) continue;
        b->batched[i] = 1;
        BatchVox& X = hv[i];
        X.g = v->g;
        X.lv = vp[i].lv;
        X.vd = v->vd;
        X.cv = cp[i].cv;
        X.boards = v->d_boards;
        X.S = v->S;
        X.chord = v->mode == CR_MODE_RGBA ? 1 : 0;
        X.lbox = vp[i].ls.lbox;
        X.counts = vp[i].ls.counts;
        X.offsets = vp[i].ls.offsets;
        X.items = vp[i].ls.items;
        X.items_cap = (int)std::min<size_t>(vp[i].ls.items_cap, 0x7fffffff);
        X.keys = v->d_keys;
        X.shell = v->d_shell_s[0];
        X.fill = v->d_fill_s[0];
        X.levels = v->d_lset[0];
        X.zflags = v->d_zflags[0];
        X.sph_world = v->d_sph_world;
        X.sph_view = v->d_sph_view;
        X.sph_tiles = v->d_sph_tiles;
        X.sortkeys = v->d_sort;
        X.order = v->d_order;
        X.tie = v->d_tie;
        X.order_mode = v->order_mode;
        X.masks = v->d_masks;
        X.mask_words = cp[i].words;
        batch_vox_complete(X);  // tile_count: once d_counts is sized below
    }
    if ((rc = ensure(ctx, &b->d_counts, (size_t)total + 2, &b->counts_cap))) return rc;
    for (size_t i = 0; i < n; ++i)
        if (b->batched[i]) hv[i].tile_count = b->d_counts + tile0[i];
    std::vector<int> ints;
    b->launches.clear();
    for (int st = 0; st < kBsCount; ++st) {
        // group the scenes by launch variant, in scene order
        std::vector<std::pair<int, std::pair<int, int>>> mem;  // (variant, (scene, blocks))
        for (size_t i = 0; i < n; ++i) {
            if (!b->batched[i]) continue;
            int blocks = 0, var = 0;
            batch_stage_shape(st, hv[i], &blocks, &var);
            if (blocks > 0) mem.push_back({var, {(int)i, blocks}});
        }
        std::stable_sort(mem.begin(), mem.end(),
                         [](const auto& a, const auto& c) { return a.first < c.first; });
        for (size_t a = 0; a < mem.size();) {
            size_t e = a;
            while (e < mem.size() && mem[e].first == mem[a].first) ++e;
            cr_batch_s::StageLaunch L{st, mem[a].first, (int)ints.size(), (int)(e - a), 0};
            int acc = 0;
            for (size_t k = a; k < e; ++k) { ints.push_back(acc); acc += mem[k].second.second; }
            ints.push_back(acc);
            for (size_t k = a; k < e; ++k) ints.push_back(mem[k].second.first);
            L.blocks = acc;
            b->launches.push_back(L);
            a = e;
        }
    }
    // the fills: depth keys, occupancy words and tile masks of the batched
    // scenes, and the tile counts of every scene
    std::vector<FillSeg> segs;
    for (size_t i = 0; i < n; ++i) {
        if (!b->batched[i]) continue;
        cr_volume v = b->items[i].volume;
        segs.push_back({(uint64_t*)v->d_keys, (uint64_t)vp[i].npx, ~0ull});
        segs.push_back({(uint64_t*)v->d_shell_s[0], (uint64_t)v->nbricks0, 0ull});
        segs.push_back({(uint64_t*)v->d_fill_s[0], (uint64_t)v->nbricks0, 0ull});
        segs.push_back({(uint64_t*)v->d_masks, (uint64_t)cp[i].nmask, 0ull});
    }
    segs.push_back({(uint64_t*)b->d_counts, (uint64_t)(total + 1) / 2, 0ull});
    b->fill_off = (int)ints.size();
    b->nseg = (int)segs.size();
    {
        int acc = 0;
        for (const FillSeg& f : segs) { ints.push_back(acc); acc += fill_segment_blocks(f.n); }
        ints.push_back(acc);
        for (int k = 0; k < b->nseg; ++k) ints.push_back(k);
        b->fill_blocks = acc;
    }
    if ((rc = ensure(ctx, &b->d_vox, n, &b->vox_cap))) return rc;
    if ((rc = ensure(ctx, &b->d_stage, ints.size(), &b->stage_cap))) return rc;
    if ((rc = ensure(ctx, &b->d_segs, segs.size(), &b->segs_cap))) return rc;
    CR_HIP(ctx, hipMemcpy(b->d_vox, hv.data(), n * sizeof(BatchVox), hipMemcpyHostToDevice));
    CR_HIP(ctx, hipMemcpy(b->d_stage, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
    CR_HIP(ctx, hipMemcpy(b->d_segs, segs.data(), segs.size() * sizeof(FillSeg), hipMemcpyHostToDevice));
    if ((rc = ensure(ctx, &b->d_scenes, n, &b->scenes_cap))) return rc;
    if ((rc = ensure(ctx, &b->d_tile0, n, &b->tile0_cap))) return rc;
    if ((rc = ensure(ctx, &b->d_order, (size_t)total, &b->order_cap))) return rc;
    if ((rc = ensure(ctx, &b->d_oscratch, 64 + (size_t)total, &b->oscratch_cap))) return rc;
    if (!b->d_counters) {
        size_t cap = 0;
        if ((rc = ensure(ctx, &b->d_counters, 16, &cap))) return rc;
    }
    CR_HIP(ctx, hipMemcpy(b->d_scenes, hs.data(), n * sizeof(BatchScene), hipMemcpyHostToDevice));
    CR_HIP(ctx, hipMemcpy(b->d_tile0, tile0.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return CR_OK;
}

// kernels only: voxelize every scene, camera setup / sort / bin every scene,
// one heavy-first order over all tiles, one batched cone trace
// kernels only.  Scenes the batched stages take: one segmented fill, then
// one launch per voxelization stage and per camera-pass stage over all of
// them; the others: their own per-scene launches.
void batch_enqueue(cr_batch b, const std::vector<VoxPlan>& vp, const std::vector<ConePlan>& cp,
                   unsigned long long* counters, hipEvent_t* ev, hipStream_t s) {
    const size_t n = b->items.size();
    if (ev) (void)hipEventRecord(ev[0], s);
    if (counters) launch_fill_u64((uint64_t*)counters, 9, 0, s);
    launch_fill_segments(b->d_segs, b->d_stage + b->fill_off, b->nseg, b->fill_blocks, s);  // tile counts included
    auto stages = [&](bool camera) {
        for (const auto& L : b->launches) {
            if ((L.stage >= kBsCam) != camera) continue;
            const BatchStage st{b->d_stage + L.first_off, b->d_stage + L.first_off + L.n + 1, L.n};
            launch_batch_stage(L.stage, L.variant, b->d_vox, st, L.blocks, s);
        }
    };
    stages(false);
    for (size_t i = 0; i < n; ++i) {
        if (b->batched[i]) continue;
        cr_volume v = b->items[i].volume;
        voxelize_enqueue(v, vp[i], nullptr, nullptr, s, nullptr, s, s, 0, nullptr);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    stages(true);
    int t0 = 0;
    for (size_t i = 0; i < n; ++i) {
        cr_volume v = b->items[i].volume;
        if (!b->batched[i]) {
            cone_enqueue_sort(v, cp[i], nullptr, s);
            launch_fill_u64((uint64_t*)v->d_masks, cp[i].nmask, 0, s);
            launch_bin_counts(v->d_sph_tiles, v->d_order, cp[i].S, cp[i].cv.tiles_x, cp[i].words, v->d_masks,
                              b->d_counts + t0, s);
        }
        t0 += (int)cp[i].ntiles;
    }
    launch_tile_order(b->d_counts, b->total_tiles, b->d_oscratch, b->d_order, s);
    if (ev) (void)hipEventRecord(ev[2], s);
    launch_cone_batch(b->d_scenes, b->d_tile0, (int)n, b->d_order, b->total_tiles, b->format, b->cone_mode,
                      b->composite, b->cfg.waves_per_group, b->cfg.groups, counters, s);
    if (ev) (void)hipEventRecord(ev[3], s);
}

// the camera buffers a batch graph bakes in (a single trace may reallocate them)
constexpr int kGraphPtrs = 12;
static void cam_ptrs(cr_volume v, const void** out) {
    const void* p[kGraphPtrs] = {v->d_sph_world, v->d_sph_view, v->d_sph_tiles, v->d_sort,
                                 v->d_order,     v->d_tie,      v->d_tiles,     v->d_masks,
                                 v->d_lbox,      v->d_lcounts,  v->d_litems,    v->d_keys};
    for (int k = 0; k < kGraphPtrs; ++k) out[k] = p[k];
}

bool graph_current(cr_batch b) {
    if (!b->have_graph || !same_cfg(b->cfg, b->graph_cfg)) return false;
    for (size_t i = 0; i < b->items.size(); ++i) {
        if (b->items[i].volume->gen != b->graph_gen[i]) return false;
        const void* p[kGraphPtrs];
        cam_ptrs(b->items[i].volume, p);
        for (int k = 0; k < kGraphPtrs; ++k)
            if (p[k] != b->graph_ptrs[kGraphPtrs * i + k]) return false;
    }
    return true;
}

int batch_capture(cr_batch b) {
    cr_context ctx = b->ctx;
    // a run of the previous graph may still read the descriptors and tables
    // batch_prepare rewrites (captures are rare: new configuration, billboards
    // or buffers)
    CR_HIP(ctx, hipDeviceSynchronize());
    drop_graph(b);
    std::vector<VoxPlan> vp;
    std::vector<ConePlan> cp;
    int rc = batch_prepare(b, vp, cp);
    if (rc) return rc;
    CR_HIP(ctx, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    batch_enqueue(b, vp, cp, nullptr, nullptr, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (e != hipSuccess) return fail(ctx, CR_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return fail(ctx, CR_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->have_graph = true;
    b->graph_cfg = b->cfg;
    b->graph_gen.resize(b->items.size());
    b->graph_ptrs.resize(kGraphPtrs * b->items.size());
    for (size_t i = 0; i < b->items.size(); ++i) {
        b->graph_gen[i] = b->items[i].volume->gen;
        cam_ptrs(b->items[i].volume, &b->graph_ptrs[kGraphPtrs * i]);
    }
    return CR_OK;
}

}  // namespace

extern "C" {

int cr_batch_create(cr_context ctx, const cr_batch_item* items, int count, cr_batch* out) {
    if (!ctx || !out || !items) return fail(ctx, CR_EINVAL, "cr_batch_create: NULL argument");
    *out = nullptr;
    if (count < 1 || count > 4096) return fail(ctx, CR_EINVAL, "cr_batch_create: count must be in [1, 4096]");
    for (int i = 0; i < count; ++i) {
        const cr_batch_item& it = items[i];
        if (!it.volume || it.volume->ctx != ctx) return fail(ctx, CR_EINVAL, "cr_batch_create: volume of another context or NULL");
        if (it.volume->format == CR_FORMAT_R8) return fail(ctx, CR_EUNSUPPORTED, "cr_batch_create: R8 volumes are not batched");
        if (it.volume->nm_on || it.volume->sun_on)
            return fail(ctx, CR_EUNSUPPORTED, "cr_batch_create: the noise march and the sun sprite are not batched");
        if (it.volume->format != items[0].volume->format || it.params.cone_mode != items[0].params.cone_mode ||
            it.params.composite != items[0].params.composite)
            return fail(ctx, CR_EUNSUPPORTED, "cr_batch_create: scenes must share format, cone mode and composite");
        for (int j = 0; j < i; ++j)
            if (items[j].volume == it.volume) return fail(ctx, CR_EINVAL, "cr_batch_create: a volume appears twice");
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    cr_batch b = new (std::nothrow) cr_batch_s();
    if (!b) return fail(ctx, CR_ENOMEM, "cr_batch_create: host allocation failed");
    b->ctx = ctx;
    b->items.assign(items, items + count);
    b->format = items[0].volume->format;
    b->cone_mode = items[0].params.cone_mode;
    b->composite = items[0].params.composite;
    hipError_t e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(ctx, CR_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    // validate every scene once (and allocate its buffers)
    std::vector<VoxPlan> vp;
    std::vector<ConePlan> cp;
    if ((rc = batch_prepare(b, vp, cp))) {
        cr_batch_destroy(b);
        return rc;
    }
    *out = b;
    return CR_OK;
}

int cr_batch_destroy(cr_batch b) {
    if (!b) return CR_OK;
    (void)hipSetDevice(b->ctx->device);
    drop_graph(b);
    void* ptrs[] = {b->d_scenes, b->d_tile0, b->d_counts, b->d_order, b->d_oscratch, b->d_counters,
                    b->d_vox,    b->d_stage, b->d_segs};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
    return CR_OK;
}

int cr_batch_set_config(cr_batch b, const cr_batch_config* cfg) {
    if (!b || !cfg) return fail(b ? b->ctx : nullptr, CR_EINVAL, "cr_batch_set_config: NULL argument");
    if (cfg->tile_w != 4 && cfg->tile_w != 8 && cfg->tile_w != 16 && cfg->tile_w != 32)
        return fail(b->ctx, CR_EINVAL, "cr_batch_set_config: tile_w must be 4, 8, 16 or 32");
    if (cfg->waves_per_group != 1 && cfg->waves_per_group != 2 && cfg->waves_per_group != 4)
        return fail(b->ctx, CR_EINVAL, "cr_batch_set_config: waves_per_group must be 1, 2 or 4");
    if (cfg->groups < 0 || cfg->groups > (1 << 20)) return fail(b->ctx, CR_EINVAL, "cr_batch_set_config: bad group count");
    b->cfg = *cfg;
    return CR_OK;
}

int cr_batch_get_config(cr_batch b, cr_batch_config* cfg) {
    if (!b || !cfg) return fail(b ? b->ctx : nullptr, CR_EINVAL, "cr_batch_get_config: NULL argument");
    *cfg = b->cfg;
    return CR_OK;
}

int cr_batch_run(cr_batch b, void* stream) {
    if (!b) return fail(nullptr, CR_EINVAL, "cr_batch_run: batch is NULL");
    cr_context ctx = b->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    for (const auto& it : b->items)
        if (it.volume->nm_on || it.volume->sun_on)
            return fail(ctx, CR_EUNSUPPORTED, "cr_batch_run: the noise march and the sun sprite are not batched");
    // the scenes' buffers are shared with their single voxelizations and
    // cone traces: wait for whatever is still queued on the context's side
    // streams (every volume's side work runs there)
    for (int k = 0; k < 3; ++k) {
        hipStream_t st = k == 0 ? ctx->lside : k == 1 ? ctx->vside : ctx->bside;
        CR_HIP(ctx, hipEventRecord(ctx->ev_side[k], st));
        CR_HIP(ctx, hipStreamWaitEvent(s, ctx->ev_side[k], 0));
    }
    for (const auto& it : b->items) {
        cr_volume v = it.volume;
        if (v->cset != 0) swap_cam(v);  // batches use camera pair 0
        if (v->lset_i != 0) swap_light(v);  // and work-list pair 0
    }
    if (ctx->profiling == 1) {
        // eager run with counters and phase times
        std::vector<VoxPlan> vp;
        std::vector<ConePlan> cp;
        CR_HIP(ctx, hipDeviceSynchronize());  // earlier runs may still read what batch_prepare rewrites
        if ((rc = batch_prepare(b, vp, cp))) return rc;
        hipEvent_t ev[4];
        for (int i = 0; i < 4; ++i) CR_HIP(ctx, hipEventCreate(&ev[i]));
        batch_enqueue(b, vp, cp, b->d_counters, ev, s);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        float t[3] = {0, 0, 0};
        for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventElapsedTime(&t[i], ev[i], ev[i + 1]);
        unsigned long long c[9] = {};
        if (e == hipSuccess) e = hipMemcpy(c, b->d_counters, sizeof(c), hipMemcpyDeviceToHost);
        for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
        if (e != hipSuccess) return fail(ctx, CR_EDEVICE, "cr_batch_run: %s", hipGetErrorString(e));
        ctx->stats = cr_stats{};
        ctx->stats.fragments = (int64_t)c[3];
        ctx->stats.samples = (int64_t)c[4];
        ctx->stats.two_level = (int64_t)c[5];
        ctx->stats.ms_setup = t[0];
        ctx->stats.ms_bin = t[1];
        ctx->stats.ms_cone = t[2];
        ctx->stats.n_cone = 1;
    } else {
        if (!graph_current(b) && (rc = batch_capture(b))) return rc;
        CR_HIP(ctx, hipGraphLaunch(b->exec, s));
    }
    // and the next camera pass and voxelization of a volume wait for this run
    for (const auto& it : b->items) {
        cr_volume v = it.volume;
        v->cam_sync_needed = true;  // its next camera pass follows this run
        v->s_sync_needed = true;  // its next voxelization follows this run
        if (v->cur != 0) {  // the batch's sets (0) become current
            CR_HIP(ctx, hipEventRecord(v->ev_release[v->cur], s));
            v->release_valid[v->cur] = true;
            v->cur = 0;
            v->d_levels = v->d_lset[0];
            v->d_shell = v->d_shell_s[0];
            v->d_fill = v->d_fill_s[0];
        }
    }
    for (auto& it : b->items) {
        it.volume->voxelized = true;
        it.volume->sorted_count = it.volume->S;
    }
    return CR_OK;
}

int cr_batch_autotune(cr_batch b, int repeats, void* stream, cr_batch_config* best, float* best_ms) {
    if (!b) return fail(nullptr, CR_EINVAL, "cr_batch_autotune: batch is NULL");
    cr_context ctx = b->ctx;
    if (repeats < 1 || repeats > 1000) return fail(ctx, CR_EINVAL, "cr_batch_autotune: repeats must be in [1, 1000]");
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    int saved_prof = ctx->profiling;
    ctx->profiling = 0;
    hipEvent_t e0, e1;
    CR_HIP(ctx, hipEventCreate(&e0));
    CR_HIP(ctx, hipEventCreate(&e1));
    const int tws[3] = {8, 16, 4};
    const int wpgs[3] = {1, 2, 4};
    const int groups[2] = {0, 2048};
    cr_batch_config bestc = b->cfg;
    float bestt = 1e30f;
    for (int a = 0; a < 3 && rc == CR_OK; ++a)
        for (int w = 0; w < 3 && rc == CR_OK; ++w)
            for (int g = 0; g < 2 && rc == CR_OK; ++g) {
                b->cfg = cr_batch_config{tws[a], wpgs[w], groups[g] / wpgs[w]};
                if ((rc = cr_batch_run(b, s))) break;  // capture + warm-up
                hipError_t e = hipEventRecord(e0, s);
                for (int r = 0; r < repeats && e == hipSuccess; ++r) e = hipGraphLaunch(b->exec, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) { rc = fail(ctx, CR_EDEVICE, "cr_batch_autotune: %s", hipGetErrorString(e)); break; }
                ms /= (float)repeats;
                if (ms < bestt) { bestt = ms; bestc = b->cfg; }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->profiling = saved_prof;
    if (rc) return rc;
    b->cfg = bestc;
    if (best) *best = bestc;
    if (best_ms) *best_ms = bestt;
    return CR_OK;
}

}  // extern "C"
