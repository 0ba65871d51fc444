// film.cpp -- Film.PostProcess and the byte conversion of Film.Save
// (Base/Film.fs:21-30, 55-66), host side.
#include <cmath>
#include <cstdint>

#include "host_scene.hpp"

namespace bnh {

// One component of Film.PostProcess: the tone curve (Vector3 arithmetic, no
// fused operations), then Vector3.Clamp(x, 0, 1) = Min(Max(x, 0), 1), whose
// .NET 9 Min/Max propagate NaN.
static float tone(float x, int tm) {
    if (tm == 1) {                                     // Aces (Film.fs:24-26)
        const float num = x * (2.51f * x + 0.03f);
        const float den = x * (2.43f * x + 0.59f) + 0.14f;
        x = num / den;
    } else if (tm == 2) {                              // Gamma (Film.fs:27-29)
        const float g = 1.0f / 2.2f;
        x = (float)std::pow((double)x, (double)g);     // MathF.Pow
    }
    if (x != x) return x;
    x = x > 0.0f ? x : 0.0f;
    return x < 1.0f ? x : 1.0f;
}

// Rgba32(Vector3) of Film.Save: v * 255 + 0.5 clamped to [0, 255], then a
// saturating float -> byte conversion (.NET 9: NaN -> 0).
static uint8_t to_byte(float v) {
    float b = v * 255.0f + 0.5f;
    if (b != b) return 0;
    b = b > 0.0f ? b : 0.0f;
    b = b < 255.0f ? b : 255.0f;
    return (uint8_t)b;
}

void postprocess(const float* film, int64_t n, int tone_mapping, float* out_rgb, uint8_t* out_rgba8) {
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = tone(film[3 * i + k], tone_mapping);
        if (out_rgb)
            for (int k = 0; k < 3; ++k) out_rgb[3 * i + k] = c[k];
        if (out_rgba8) {
            for (int k = 0; k < 3; ++k) out_rgba8[4 * i + k] = to_byte(c[k]);
            out_rgba8[4 * i + 3] = 255;
        }
    }
}

}  // namespace bnh
