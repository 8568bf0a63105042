// pt_hdr.cpp -- Radiance RGBE (.hdr) reader for environment maps (host C++).
//
// The reference loads environment images with stb_image's stbi_loadf (EnvironmentMap.cpp:191-210,
// a third-party dependency the reference does not vendor). This is the library's own reader of the
// published Radiance picture format, restricted to what that call path accepts for .hdr files:
//   - a "#?RADIANCE" or "#?RGBE" signature line, header lines up to a blank line, where a FORMAT
//     line must say 32-bit_rle_rgbe;
//   - the standard orientation line "-Y <height> +X <width>" (rows top to bottom);
//   - each scanline either flat (4 bytes per pixel) or, for widths 8..32767, in the adaptive
//     run-length form: the marker 2 2 (width >> 8) (width & 255), then the four channels one after
//     another, each as runs (count > 128: one byte repeated count - 128 times) and literals
//     (count <= 128: that many bytes);
//   - texel = (r, g, b) * 2^(e - 136), and 0 when e = 0.
#include "../../include/pt_hip.h"
#include "pt_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

struct File {
    std::FILE* f = nullptr;
    ~File()
    {
        if (f != nullptr) {
            std::fclose(f);
        }
    }
};

bool read_line(std::FILE* f, std::string& line)
{
    line.clear();
    for (;;) {
        const int c = std::fgetc(f);
        if (c == EOF) {
            return !line.empty();
        }
        if (c == '\n') {
            return true;
        }
        line.push_back(static_cast<char>(c));
    }
}

float texel_channel(unsigned char m, unsigned char e)
{
    return static_cast<float>(m) * std::ldexp(1.0f, static_cast<int>(e) - 136);
}

// One scanline of `width` RGBE pixels into out (4 bytes per pixel).
bool read_scanline(std::FILE* f, int width, unsigned char* out)
{
    unsigned char head[4];
    if (std::fread(head, 1, 4, f) != 4) {
        return false;
    }
    const bool rle = width >= 8 && width < 32768 && head[0] == 2 && head[1] == 2 && (head[2] & 0x80) == 0;
    if (!rle) {
        std::memcpy(out, head, 4);
        return width <= 1 || std::fread(out + 4, 4, static_cast<size_t>(width - 1), f) == static_cast<size_t>(width - 1);
    }
    if (((head[2] << 8) | head[3]) != width) {
        return false;
    }
    for (int ch = 0; ch < 4; ++ch) {
        int x = 0;
        while (x < width) {
            const int count = std::fgetc(f);
            if (count == EOF || count == 0) {
                return false;
            }
            if (count > 128) {
                const int n = count - 128;
                const int v = std::fgetc(f);
                if (v == EOF || x + n > width) {
                    return false;
                }
                for (int k = 0; k < n; ++k) {
                    out[4 * (x + k) + ch] = static_cast<unsigned char>(v);
                }
                x += n;
            } else {
                if (x + count > width) {
                    return false;
                }
                for (int k = 0; k < count; ++k) {
                    const int v = std::fgetc(f);
                    if (v == EOF) {
                        return false;
                    }
                    out[4 * (x + k) + ch] = static_cast<unsigned char>(v);
                }
                x += count;
            }
        }
    }
    return true;
}

} // namespace

extern "C" {

float* ptEnvironmentLoadHdr(const char* path, int* width, int* height)
{
    if (path == nullptr || width == nullptr || height == nullptr) {
        ptSetError("ptEnvironmentLoadHdr: invalid arguments");
        return nullptr;
    }
    File file;
    file.f = std::fopen(path, "rb");
    if (file.f == nullptr) {
        ptSetError("ptEnvironmentLoadHdr: cannot open %s", path);
        return nullptr;
    }
    std::string line;
    if (!read_line(file.f, line) || (line != "#?RADIANCE" && line != "#?RGBE")) {
        ptSetError("ptEnvironmentLoadHdr: %s is not a Radiance picture", path);
        return nullptr;
    }
    bool formatOk = false;
    while (read_line(file.f, line) && !line.empty()) {
        if (line.rfind("FORMAT=", 0) == 0) {
            formatOk = line == "FORMAT=32-bit_rle_rgbe";
            if (!formatOk) {
                ptSetError("ptEnvironmentLoadHdr: unsupported %s", line.c_str());
                return nullptr;
            }
        }
    }
    if (!formatOk) {
        ptSetError("ptEnvironmentLoadHdr: missing FORMAT=32-bit_rle_rgbe");
        return nullptr;
    }
    int h = 0, w = 0;
    char ya[3] = {}, xa[3] = {};
    if (!read_line(file.f, line) || std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 ||
        std::strcmp(ya, "-Y") != 0 || std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0 || w > (1 << 24) / h) {
        ptSetError("ptEnvironmentLoadHdr: unsupported resolution line \"%s\" (expected -Y <h> +X <w>)", line.c_str());
        return nullptr;
    }
    float* rgb = static_cast<float*>(std::malloc(sizeof(float) * 3 * static_cast<size_t>(w) * static_cast<size_t>(h)));
    if (rgb == nullptr) {
        ptSetError("ptEnvironmentLoadHdr: out of memory");
        return nullptr;
    }
    std::vector<unsigned char> row(4 * static_cast<size_t>(w));
    for (int y = 0; y < h; ++y) {
        if (!read_scanline(file.f, w, row.data())) {
            std::free(rgb);
            ptSetError("ptEnvironmentLoadHdr: truncated or malformed scanline %d of %s", y, path);
            return nullptr;
        }
        float* dst = rgb + 3 * static_cast<size_t>(y) * static_cast<size_t>(w);
        for (int x = 0; x < w; ++x) {
            const unsigned char* p = &row[4 * static_cast<size_t>(x)];
            for (int c = 0; c < 3; ++c) {
                dst[3 * x + c] = p[3] == 0 ? 0.0f : texel_channel(p[c], p[3]);
            }
        }
    }
    *width = w;
    *height = h;
    return rgb;
}

void ptEnvironmentFreeImage(float* rgb) { std::free(rgb); }

} // extern "C"
