// Image files of the façade: PNG output (uncompressed deflate blocks) and
// Radiance RGBE (.hdr) / PFM environment input.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ohao/renderer.hpp"

namespace ohao {
namespace {

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    c ^= 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put32(out, crc32(td.data(), td.size()));
}

}  // namespace

bool writePNG(const std::string& path, const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (!rgba || width == 0 || height == 0) return false;
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (width * 4 + 1));
    for (uint32_t y = 0; y < height; ++y) {
        raw.push_back(0);   // filter: none
        raw.insert(raw.end(), rgba + (size_t)y * width * 4, rgba + (size_t)(y + 1) * width * 4);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    while (pos < raw.size() || raw.empty()) {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(n & 0xFF));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xFF));
        z.push_back((uint8_t)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + n));
        pos += n;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> ihdr;
    put32(ihdr, width);
    put32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
    return (bool)f;
}

namespace {

bool read_pfm(std::ifstream& f, std::vector<float>& rgba, uint32_t& W, uint32_t& H, bool color) {
    int w = 0, h = 0;
    double scale = 0.0;
    f >> w >> h >> scale;
    f.get();
    if (w <= 0 || h <= 0) return false;
    const int c = color ? 3 : 1;
    std::vector<float> data((size_t)w * h * c);
    f.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * 4));
    if (!f) return false;
    if (scale > 0) {   // big-endian samples
        for (float& v : data) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&v, &u, 4);
        }
    }
    W = (uint32_t)w;
    H = (uint32_t)h;
    rgba.assign((size_t)w * h * 4, 1.0f);
    for (int y = 0; y < h; ++y)          // PFM rows run bottom to top
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k)
                rgba[((size_t)y * w + x) * 4 + k] = data[((size_t)(h - 1 - y) * w + x) * c + (color ? k : 0)];
    return true;
}

}  // namespace

bool readHDR(const std::string& path, std::vector<float>& rgba, uint32_t& W, uint32_t& H) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string line;
    std::getline(f, line);
    if (line == "PF" || line == "Pf") return read_pfm(f, rgba, W, H, line == "PF");
    if (line.rfind("#?", 0) != 0) return false;
    bool rgbe = false;
    while (std::getline(f, line) && !line.empty())
        if (line.find("FORMAT=32-bit_rle_rgbe") != std::string::npos) rgbe = true;
    if (!rgbe) return false;
    std::getline(f, line);
    char ya[3] = {}, xa[3] = {};
    int h = 0, w = 0;
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::strcmp(ya, "-Y") != 0 ||
        std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0)
        return false;
    std::vector<uint8_t> scan((size_t)w * 4);
    rgba.assign((size_t)w * h * 4, 1.0f);
    for (int y = 0; y < h; ++y) {
        uint8_t hdr[4];
        if (!f.read(reinterpret_cast<char*>(hdr), 4)) return false;
        if (w >= 8 && w < 32768 && hdr[0] == 2 && hdr[1] == 2 && ((hdr[2] << 8) | hdr[3]) == w) {
            for (int c = 0; c < 4; ++c) {   // run-length encoded, one component at a time
                int x = 0;
                while (x < w) {
                    int n = f.get();
                    if (n == EOF) return false;
                    if (n > 128) {
                        n -= 128;
                        const int v = f.get();
                        if (v == EOF || x + n > w) return false;
                        for (int k = 0; k < n; ++k) scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
                    } else {
                        if (n == 0 || x + n > w) return false;
                        for (int k = 0; k < n; ++k) {
                            const int v = f.get();
                            if (v == EOF) return false;
                            scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
                        }
                    }
                }
            }
        } else {   // flat RGBE pixels
            std::memcpy(scan.data(), hdr, 4);
            if (!f.read(reinterpret_cast<char*>(scan.data() + 4), (std::streamsize)(scan.size() - 4))) return false;
        }
        for (int x = 0; x < w; ++x) {
            const uint8_t* p = &scan[(size_t)x * 4];
            const float s = p[3] ? std::ldexp(1.0f, (int)p[3] - 136) : 0.0f;
            float* o = &rgba[((size_t)y * w + x) * 4];
            o[0] = p[0] * s;
            o[1] = p[1] * s;
            o[2] = p[2] * s;
        }
    }
    W = (uint32_t)w;
    H = (uint32_t)h;
    return true;
}

}  // namespace ohao
