// cloudvox.hpp - header-only C++ facade over the C ABI (cloudvox.h) that
// mirrors the host objects of Cloud-Renderer's render path, so that code
// written against the reference reads the same:
//
//   cr::CloudVolume      CloudVolume (src/CloudVolume.hpp:12-55): position,
//                        x/y/zBounds, dimension, levels, voxelSize,
//                        billboards {positions, scales}, fluffiness, update()
//   cr::Sun              Sun (src/Sun.hpp:12-19): position (+ directional sun),
//                        sprite colours and radii
//   cr::SunShader        SunShader::render() (src/Shaders/SunShader.cpp:7-42):
//                        the sprite drawn before the clouds
//   cr::Camera           Camera (src/Camera.hpp:25-45): position, lookAt, and
//                        the P/V the reference uploads (src/Camera.cpp:59-60)
//   cr::VoxelizeShader   VoxelizeShader::voxelize(CloudVolume*)
//                        (src/Shaders/VoxelizeShader.hpp:18)
//   cr::ConeTraceShader  ConeTraceShader::coneTrace(CloudVolume*) with the
//                        vct* parameters (src/Shaders/ConeTraceShader.hpp:12-32)
//
// Errors throw cr::Error (C++ only; the C ABI itself never throws).
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "cloudvox.h"

namespace cr {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, cr_context ctx) {
    if (rc != CR_OK) throw Error(rc, std::string("cloudvox: ") + cr_last_error(ctx));
}

using vec2 = std::array<float, 2>;
using vec3 = std::array<float, 3>;

class Context {
  public:
    explicit Context(int device = 0) { check(cr_context_create(device, 0, &h_), nullptr); }
    ~Context() {
        if (h_) cr_context_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cr_context handle() const { return h_; }
    void setProfiling(int level) { check(cr_context_set_profiling(h_, level), h_); }
    cr_stats stats() const {
        cr_stats s;
        check(cr_get_stats(h_, &s), h_);
        return s;
    }

  private:
    cr_context h_ = nullptr;
};

class CloudVolume {
  public:
    struct Billboards {
        std::vector<vec3> positions;
        std::vector<float> scales;
        int count() const { return (int)positions.size(); }
    };

    // CloudVolume(int dim, glm::vec2 bounds, glm::vec3 position, int mips)
    CloudVolume(Context& ctx, int dim, vec2 bounds, vec3 pos, int mips, int format = CR_FORMAT_RGBA32F,
                int mode = CR_MODE_RGBA)
        : position(pos), xBounds(bounds), yBounds(bounds), zBounds(bounds), dimension(dim), levels(mips),
          ctx_(&ctx) {
        float lo[3] = {bounds[0], bounds[0], bounds[0]};
        float hi[3] = {bounds[1], bounds[1], bounds[1]};
        check(cr_volume_create(ctx.handle(), dim, lo, hi, pos.data(), mips, format, mode, &h_), ctx.handle());
        for (int a = 0; a < 3; ++a) voxelSize[a] = (bounds[1] - bounds[0]) / (float)dim;
    }
    ~CloudVolume() {
        if (h_) cr_volume_destroy(h_);
    }
    CloudVolume(const CloudVolume&) = delete;
    CloudVolume& operator=(const CloudVolume&) = delete;

    void addCloudBoard(const vec3& p, float scale) {
        billboards.positions.push_back(p);
        billboards.scales.push_back(scale);
    }
    // CloudVolume::update -> uploadBillboards (src/CloudVolume.cpp:84-93, 139-164)
    void update() {
        std::vector<float> xyz;
        xyz.reserve(3 * billboards.positions.size());
        for (const vec3& p : billboards.positions) xyz.insert(xyz.end(), p.begin(), p.end());
        check(cr_volume_set_billboards(h_, xyz.data(), billboards.scales.data(), billboards.count(), fluffiness),
              ctx_->handle());
    }
    void setNoise(int octaves, uint32_t seed, float baseFreq = 4.0f, float lacunarity = 2.0f, float gain = 0.5f) {
        check(cr_volume_set_noise(h_, octaves, seed, baseFreq, lacunarity, gain), ctx_->handle());
    }
    // occupied voxels (shell and chord fill) plus an order-free hash
    std::pair<uint64_t, uint64_t> occupancy(void* stream = nullptr) const {
        uint64_t c = 0, h = 0;
        check(cr_volume_occupancy(h_, &c, &h, stream), ctx_->handle());
        return {c, h};
    }
    // VoxelShader's activeVoxels (src/Shaders/VoxelShader.cpp:110-115): level-0 texels with r != 0
    uint64_t activeVoxels(void* stream = nullptr) const {
        uint64_t c = 0;
        check(cr_volume_active_voxels(h_, &c, stream), ctx_->handle());
        return c;
    }
    // CR_ORDER_SORTED (traces sort first) or CR_ORDER_UPLOAD (draw in upload order, as the reference's
    // draw call uses the buffers update() uploaded)
    void setDrawOrder(int mode) { check(cr_volume_set_draw_order(h_, mode), ctx_->handle()); }
    // CloudVolume::sortBoards (src/CloudVolume.cpp:65-82): billboards permuted in place, far to near
    // from point, equal distances as the selection sort leaves them; update() uploads the new order
    void sortBoards(const vec3& point, void* stream = nullptr) {
        const int n = billboards.count();
        std::vector<int32_t> order((size_t)(n > 0 ? n : 1));
        check(cr_volume_sort_boards(h_, point.data(), order.data(), (int)order.size(), stream), ctx_->handle());
        Billboards sorted;
        for (int r = 0; r < n; ++r) {
            sorted.positions.push_back(billboards.positions[(size_t)order[(size_t)r]]);
            sorted.scales.push_back(billboards.scales[(size_t)order[(size_t)r]]);
        }
        billboards = sorted;
    }
    // level k, x-fastest, native element type (u8 / float / binary16)
    void readLevel(int level, void* dst, size_t bytes, void* stream = nullptr) const {
        check(cr_volume_read_level(h_, level, dst, bytes, stream), ctx_->handle());
    }

    cr_volume handle() const { return h_; }
    cr_context context() const { return ctx_->handle(); }

    vec3 position;
    vec2 xBounds, yBounds, zBounds;
    int dimension;
    int levels;
    vec3 voxelSize{};
    Billboards billboards;
    float fluffiness = 1.0f;

  private:
    Context* ctx_;
    cr_volume h_ = nullptr;
};

struct Sun {
    vec3 position{5.0f, 20.0f, -5.0f};  // src/main.cpp:37
    vec3 direction{0.0f, 1.0f, 0.0f};
    bool directional = false;
    // sprite (src/main.cpp:40-43)
    vec3 innerColor{1.0f, 1.0f, 1.0f};
    vec3 outerColor{1.0f, 1.0f, 0.0f};
    float innerRadius = 1.0f;
    float outerRadius = 2.0f;
    cr_light c() const {
        cr_light l{};
        for (int a = 0; a < 3; ++a) {
            l.position[a] = position[a];
            l.direction[a] = direction[a];
        }
        l.directional = directional ? 1 : 0;
        return l;
    }
};

// Camera::update (src/Camera.cpp:59-60): V = lookAt(position, lookAt, up),
// P = perspective(fovy, aspect, znear, zfar).  The reference passes fovy = 45
// (glm takes it as radians) and aspect = (float)(Window::width /
// Window::height), an integer division; aspect 0 here means width / height.
// setMatrices() takes P and V as the reference uploads them instead
// (column-major, src/Shaders/ConeTraceShader.cpp:64-69); position stays the
// point billboards are sorted from.
struct Camera {
    vec3 position{0.0f, 0.0f, 3.0f};
    vec3 lookAt{0.0f, 0.0f, 0.0f};
    vec3 up{0.0f, 1.0f, 0.0f};
    float fovy = 1.0471976f;
    float znear = 0.01f, zfar = 2500.0f;  // src/Camera.cpp:59
    float aspect = 0.0f;
    bool useMatrices = false;
    std::array<float, 16> P{}, V{};
    void setMatrices(const float* p, const float* v) {
        for (int i = 0; i < 16; ++i) {
            P[i] = p[i];
            V[i] = v[i];
        }
        useMatrices = true;
    }
    cr_camera c() const {
        cr_camera k{};
        for (int a = 0; a < 3; ++a) {
            k.eye[a] = position[a];
            k.target[a] = lookAt[a];
            k.up[a] = up[a];
        }
        k.fovy = fovy;
        k.znear = znear;
        k.zfar = zfar;
        k.aspect = aspect;
        k.use_matrices = useMatrices ? 1 : 0;
        for (int i = 0; i < 16; ++i) {
            k.P[i] = P[i];
            k.V[i] = V[i];
        }
        return k;
    }
};

// SunShader::render (src/Shaders/SunShader.cpp:7-42, drawn after the clear
// and before the clouds, src/main.cpp:112-116): the sprite is composited by
// the volume's next cone trace (GL-blend composite only).
class SunShader {
  public:
    void render(CloudVolume* volume, const Sun& sun, bool enabled = true) const {
        cr_sun_sprite q{};
        q.enabled = enabled ? 1 : 0;
        for (int a = 0; a < 3; ++a) {
            q.inner_color[a] = sun.innerColor[a];
            q.outer_color[a] = sun.outerColor[a];
        }
        q.inner_radius = sun.innerRadius;
        q.outer_radius = sun.outerRadius;
        check(cr_volume_set_sun_sprite(volume->handle(), &q), volume->context());
    }
};

class VoxelizeShader {
  public:
    VoxelizeShader(int width, int height) : width_(width), height_(height) {}
    // VoxelizeShader::voxelize (src/Shaders/VoxelizeShader.cpp:18-31)
    void voxelize(CloudVolume* volume, const Sun& sun, void* stream = nullptr) const {
        cr_light l = sun.c();
        check(cr_voxelize(volume->handle(), &l, width_, height_, stream), volume->context());
    }

  private:
    int width_, height_;
};

class ConeTraceShader {
  public:
    // src/Shaders/ConeTraceShader.hpp:28-32
    int vctSteps = 16;
    float vctConeAngle = 0.9f;
    float vctConeInitialHeight = 0.1f;
    float vctLodOffset = 0.0f;
    float vctDownScaling = 1.0f;
    int coneMode = CR_CONE_WEIGHTED;
    float alphaStop = 0.99f;
    float sigma = 24.0f;
    int composite = CR_COMPOSITE_GL_BLEND;
    bool quantizeRGBA8 = true;
    bool compensated = false;
    std::array<float, 4> clearColor{0.2f, 0.3f, 0.5f, 1.0f};  // src/main.cpp:112
    // noise march (src/Shaders/ConeTraceShader.hpp:14-25, 35-36); the map's
    // alpha bytes come from noiseSeed in place of rand()
    bool doNoiseSample = true;
    uint32_t noiseSeed = 0;
    float stepSize = 0.01f;
    float noiseOpacity = 4.0f;
    int numOctaves = 4;
    float freqStep = 3.0f;
    float persStep = 0.5f;
    float adjustSize = 40.0f;
    int minNoiseSteps = 2;
    int maxNoiseSteps = 8;
    float minNoiseColor = 0.2f;
    float noiseColorScale = 0.45f;
    vec3 windVel{0.01f, 0.0f, 0.0f};
    float runTime = 0.0f;  // Window::runTime: octaveOffsets = windVel * runTime

    cr_noise_march noiseParams() const {
        cr_noise_march m{};
        m.enabled = doNoiseSample ? 1 : 0;
        m.seed = noiseSeed;
        m.step_size = stepSize;
        m.opacity = noiseOpacity;
        m.octaves = numOctaves;
        m.freq_step = freqStep;
        m.pers_step = persStep;
        m.adjust_size = adjustSize;
        m.min_steps = minNoiseSteps;
        m.max_steps = maxNoiseSteps;
        m.min_color = minNoiseColor;
        m.color_scale = noiseColorScale;
        for (int a = 0; a < 3; ++a) m.offsets[a] = windVel[a] * runTime;
        return m;
    }

    cr_cone_params params() const {
        cr_cone_params p{};
        p.steps = vctSteps;
        p.cone_angle = vctConeAngle;
        p.initial_height = vctConeInitialHeight;
        p.lod_offset = vctLodOffset;
        p.down_scaling = vctDownScaling;
        p.cone_mode = coneMode;
        p.alpha_stop = alphaStop;
        p.sigma = sigma;
        p.composite = composite;
        p.quantize_rgba8 = quantizeRGBA8 ? 1 : 0;
        p.compensated = compensated ? 1 : 0;
        for (int q = 0; q < 4; ++q) p.clear_color[q] = clearColor[q];
        return p;
    }

    // ConeTraceShader::coneTrace (src/Shaders/ConeTraceShader.cpp:15-82) into a device image
    void coneTrace(CloudVolume* volume, const Sun& sun, const Camera& camera, void* d_image, int width, int height,
                   size_t pitch, int pixelFormat = CR_PIXEL_RGBA32F, void* stream = nullptr) const {
        cr_cone_params p = params();
        cr_light l = sun.c();
        cr_camera k = camera.c();
        cr_noise_march m = noiseParams();
        check(cr_volume_set_noise_march(volume->handle(), &m), volume->context());
        check(cr_cone_trace(volume->handle(), &l, &k, &p, d_image, width, height, pitch, pixelFormat, stream),
              volume->context());
    }

    // coneTrace with the far field after k samples predicted by `mlp`
    void coneTraceLearned(CloudVolume* volume, const Sun& sun, const Camera& camera, cr_mlp mlp, int k,
                          float errBound, void* d_image, int width, int height, size_t pitch,
                          int pixelFormat = CR_PIXEL_RGBA32F, void* stream = nullptr) const {
        cr_cone_params p = params();
        cr_light l = sun.c();
        cr_camera c = camera.c();
        check(cr_volume_set_noise_march(volume->handle(), nullptr), volume->context());
        check(cr_cone_trace_learned(volume->handle(), &l, &c, &p, mlp, k, errBound, d_image, width, height, pitch,
                                    pixelFormat, stream),
              volume->context());
    }
};

// The learned far field (cr_mlp): parameters in the CR_MLP_PARAMS layout.
class FarFieldMLP {
  public:
    explicit FarFieldMLP(Context& ctx) : ctx_(ctx.handle()) { check(cr_mlp_create(ctx_, &h_), ctx_); }
    FarFieldMLP(const FarFieldMLP&) = delete;
    FarFieldMLP& operator=(const FarFieldMLP&) = delete;
    ~FarFieldMLP() { cr_mlp_destroy(h_); }
    void setParams(const std::vector<float>& p) { check(cr_mlp_set_params(h_, p.data(), (int)p.size()), ctx_); }
    std::vector<float> params() const {
        std::vector<float> p(CR_MLP_PARAMS);
        check(cr_mlp_get_params(h_, p.data(), (int)p.size()), ctx_);
        return p;
    }
    cr_mlp handle() const { return h_; }

  private:
    cr_context ctx_;
    cr_mlp h_ = nullptr;
};

// Many scenes in one hipGraph (cr_batch).
class Batch {
  public:
    Batch(Context& ctx, const std::vector<cr_batch_item>& items) : ctx_(ctx.handle()) {
        check(cr_batch_create(ctx_, items.data(), (int)items.size(), &h_), ctx_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    ~Batch() { cr_batch_destroy(h_); }
    void run(void* stream = nullptr) { check(cr_batch_run(h_, stream), ctx_); }
    cr_batch_config autotune(int repeats = 3, void* stream = nullptr) {
        cr_batch_config best{};
        float ms = 0.0f;
        check(cr_batch_autotune(h_, repeats, stream, &best, &ms), ctx_);
        return best;
    }

  private:
    cr_context ctx_;
    cr_batch h_ = nullptr;
};

}  // namespace cr
