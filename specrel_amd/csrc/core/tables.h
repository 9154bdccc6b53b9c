// Lookup tables uploaded beside every scene: the 4096-entry gamma LUT (renderer.rs:1102-1114)
// and the 5x7 glyph rows of the clock readout (hud.rs:473-532 for '0'-'9', '+', '-', '.', 'S',
// in the order of gpu.rs:33-35). Host-only.
#pragma once
#include <stdint.h>
#include "minkowski.h"

namespace sr {

static const char kShaderGlyphs[14] = {'0', '1', '2', '3', '4', '5', '6',
                                       '7', '8', '9', '+', '-', '.', 'S'};

inline void glyph_5x7(char ch, uint8_t out[7]) {
    static const uint8_t digits[10][7] = {
        {0x0E, 0x11, 0x13, 0x15, 0x19, 0x11, 0x0E}, {0x04, 0x0C, 0x04, 0x04, 0x04, 0x04, 0x0E},
        {0x0E, 0x11, 0x01, 0x02, 0x04, 0x08, 0x1F}, {0x1E, 0x01, 0x01, 0x0E, 0x01, 0x01, 0x1E},
        {0x02, 0x06, 0x0A, 0x12, 0x1F, 0x02, 0x02}, {0x1F, 0x10, 0x1E, 0x01, 0x01, 0x11, 0x0E},
        {0x0E, 0x10, 0x10, 0x1E, 0x11, 0x11, 0x0E}, {0x1F, 0x01, 0x02, 0x04, 0x08, 0x08, 0x08},
        {0x0E, 0x11, 0x11, 0x0E, 0x11, 0x11, 0x0E}, {0x0E, 0x11, 0x11, 0x0F, 0x01, 0x01, 0x0E}};
    static const uint8_t plus[7] = {0x00, 0x04, 0x04, 0x1F, 0x04, 0x04, 0x00};
    static const uint8_t minus[7] = {0x00, 0x00, 0x00, 0x1F, 0x00, 0x00, 0x00};
    static const uint8_t dot[7] = {0x00, 0x00, 0x00, 0x00, 0x00, 0x0C, 0x0C};
    static const uint8_t ess[7] = {0x0F, 0x10, 0x10, 0x0E, 0x01, 0x01, 0x1E};
    const uint8_t *src = nullptr;
    if (ch >= '0' && ch <= '9') src = digits[ch - '0'];
    else if (ch == '+') src = plus;
    else if (ch == '-') src = minus;
    else if (ch == '.') src = dot;
    else if (ch == 'S') src = ess;
    for (int i = 0; i < 7; ++i) out[i] = src ? src[i] : 0;
}

// gamma_lut: (x^(1/2.2) * 255 + 0.5).clamp(0, 255) as u8, x = i / 4095.
inline void build_tables(uint32_t *out /* 4096 + 98 words */) {
    const float inv_gamma = 1.0f / 2.2f;
    for (int i = 0; i < 4096; ++i) {
        float x = (float)i / 4095.0f;
        float v = clamp_(powf_(x, inv_gamma) * 255.0f + 0.5f, 0.0f, 255.0f);
        out[i] = (uint32_t)(uint8_t)f2u_sat(v);
    }
    for (int g = 0; g < 14; ++g) {
        uint8_t rows[7];
        glyph_5x7(kShaderGlyphs[g], rows);
        for (int r = 0; r < 7; ++r) out[4096 + g * 7 + r] = rows[r];
    }
}

}  // namespace sr
