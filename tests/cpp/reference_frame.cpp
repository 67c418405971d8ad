// The reference application's default frame rendered through the C++ facade
// (include/cloudvox.hpp) and libcloudvox.so, written the way
// src/main.cpp:69-124 drives the GL passes: a CloudVolume(32, (-5, 5),
// (25, 0, 0), 4) holding GL_R8, its billboards uploaded by update(), the sun
// at (5, 20, -5), the start-up camera with P = perspective(45.f,
// (float)(width / height), .01f, 2500.f), the sun sprite, voxelize() and
// coneTrace() with the ConeTraceShader defaults (noise march on) into an
// RGBA8 image cleared to (0.2, 0.3, 0.5, 1).
//
//   reference_frame <billboards.bin> <width> <height> <noise seed> <out.bin>
//
// billboards.bin: uint32 count, then count * 3 float offsets and count float
// scales.  out.bin: uint64 activeVoxels, uint64 occupied voxels, uint64
// occupancy hash, then height * width * 4 bytes (bottom row first).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cloudvox.hpp"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s billboards.bin width height seed out.bin\n", argv[0]);
        return 2;
    }
    const int width = std::atoi(argv[2]), height = std::atoi(argv[3]);
    const uint32_t seed = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t count = 0;
    if (std::fread(&count, 4, 1, f) != 1) return 3;
    std::vector<float> xyz(3 * (size_t)count), scales(count);
    if (std::fread(xyz.data(), 4, xyz.size(), f) != xyz.size() ||
        std::fread(scales.data(), 4, scales.size(), f) != scales.size())
        return 3;
    std::fclose(f);

    try {
        cr::Context ctx(0);
        // src/main.cpp:29-34, 81-82
        cr::CloudVolume volume(ctx, 32, {-5.0f, 5.0f}, {25.0f, 0.0f, 0.0f}, 4, CR_FORMAT_R8, CR_MODE_REF);
        for (uint32_t i = 0; i < count; ++i)
            volume.addCloudBoard({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, scales[i]);
        cr::Sun sun;  // position (5, 20, -5), sprite defaults (src/main.cpp:37-43)
        // Camera::update with phi = theta = 0 and position 0 (src/Camera.cpp:18-60)
        cr::Camera camera;
        camera.position = {0.0f, 0.0f, 0.0f};
        const double theta = 0.0, phi = 0.0;
        const float PI = 3.14159265359f;
        cr::vec3 sphere{(float)(std::cos(phi) * std::cos(theta)), (float)std::sin(phi),
                        (float)(std::cos(phi) * std::cos((PI / 2.f) - theta))};
        float n = 1.0f / std::sqrt(sphere[0] * sphere[0] + sphere[1] * sphere[1] + sphere[2] * sphere[2]);
        camera.lookAt = {sphere[0] * n, sphere[1] * n, sphere[2] * n};
        camera.fovy = 45.f;
        camera.aspect = (float)(width / height);
        camera.znear = 0.01f;
        camera.zfar = 2500.f;

        cr::VoxelizeShader voxelizeShader(width, height);
        cr::SunShader sunShader;
        cr::ConeTraceShader coneShader;  // vctSteps 16, angle 0.9, ... doNoiseSample on
        coneShader.noiseSeed = seed;

        uint8_t* d_image = nullptr;
        if (hipMalloc((void**)&d_image, (size_t)width * height * 4) != hipSuccess) return 4;

        // one frame (src/main.cpp:98-124)
        volume.update();
        sunShader.render(&volume, sun);
        voxelizeShader.voxelize(&volume, sun);
        coneShader.coneTrace(&volume, sun, camera, d_image, width, height, (size_t)width * 4, CR_PIXEL_RGBA8);
        if (hipDeviceSynchronize() != hipSuccess) return 4;

        std::vector<uint8_t> image((size_t)width * height * 4);
        if (hipMemcpy(image.data(), d_image, image.size(), hipMemcpyDeviceToHost) != hipSuccess) return 4;
        if (hipFree(d_image) != hipSuccess) return 4;
        uint64_t active = volume.activeVoxels();
        std::pair<uint64_t, uint64_t> occ = volume.occupancy();

        std::FILE* o = std::fopen(argv[5], "wb");
        if (!o) return 5;
        std::fwrite(&active, 8, 1, o);
        std::fwrite(&occ.first, 8, 1, o);
        std::fwrite(&occ.second, 8, 1, o);
        std::fwrite(image.data(), 1, image.size(), o);
        std::fclose(o);
        std::printf("reference_frame: %dx%d activeVoxels %llu\n", width, height, (unsigned long long)active);
    } catch (const cr::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
