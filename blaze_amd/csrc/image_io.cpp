// image_io.cpp — frame writers of the renderer's output.
//
// Reference: RenderEngine::writeToDisk (asiroisvigneux/Blaze
// src/RenderEngine.cu:367-476): .ppm as ASCII P3 with int(255.99 * v) per
// channel, .png as 8-bit RGBA with (int)(v * 255.0) clamped to [0, 255]
// (the reference links lodepng), .exr as three FLOAT channels stored B, G, R
// (the reference links tinyexr with RLE). Written here from the file-format
// specifications: PNG (RFC 2083: signature, IHDR/IDAT/IEND chunks with
// CRC-32, a zlib stream (RFC 1950/1951) of stored deflate blocks with
// Adler-32), OpenEXR (single-part scanline file, no compression, one
// scanline per block, offset table).
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu) {
    if (!crc_ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            crc_table[i] = v;
        }
        crc_ready = true;
    }
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    be32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    be32(o, crc32(o.data() + start, o.size() - start) ^ 0xffffffffu);
}

bool write_file(const std::string& path, const std::vector<uint8_t>& bytes) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    return std::fclose(f) == 0 && ok;
}

int to8(float v) {
    int q = (int)(v * 255.0);
    return q < 0 ? 0 : (q > 255 ? 255 : q);
}

bool write_png(const std::string& path, const float* rgb, int w, int h) {
    // raw scanlines: filter byte 0, then RGBA
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (1 + 4 * (size_t)w));
    for (int j = 0; j < h; ++j) {
        raw.push_back(0);
        for (int i = 0; i < w; ++i) {
            const float* p = rgb + 3 * ((size_t)i + (size_t)w * j);
            raw.push_back((uint8_t)to8(p[0]));
            raw.push_back((uint8_t)to8(p[1]));
            raw.push_back((uint8_t)to8(p[2]));
            raw.push_back(255);
        }
    }
    // zlib: CMF/FLG, stored blocks of at most 65535 bytes, Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(n & 0xff));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xff));
        z.push_back((uint8_t)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + (std::ptrdiff_t)pos, raw.begin() + (std::ptrdiff_t)(pos + n));
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    be32(z, (b << 16) | a);
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    std::vector<uint8_t> ihdr;
    be32(ihdr, (uint32_t)w);
    be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0}); // 8-bit RGBA, deflate, adaptive filter, no interlace
    chunk(o, "IHDR", ihdr);
    chunk(o, "IDAT", z);
    chunk(o, "IEND", {});
    return write_file(path, o);
}

bool write_ppm(const std::string& path, const float* rgb, int w, int h) {
    std::string s = "P3\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    char buf[48];
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        std::snprintf(buf, sizeof(buf), "%d %d %d\n", int(255.99 * rgb[3 * i]), int(255.99 * rgb[3 * i + 1]),
                      int(255.99 * rgb[3 * i + 2]));
        s += buf;
    }
    return write_file(path, std::vector<uint8_t>(s.begin(), s.end()));
}

void le32(std::vector<uint8_t>& o, uint32_t v) {
    for (int k = 0; k < 4; ++k) o.push_back((uint8_t)(v >> (8 * k)));
}
void le64(std::vector<uint8_t>& o, uint64_t v) {
    for (int k = 0; k < 8; ++k) o.push_back((uint8_t)(v >> (8 * k)));
}
void attr(std::vector<uint8_t>& o, const char* name, const char* type, const std::vector<uint8_t>& v) {
    o.insert(o.end(), name, name + std::strlen(name) + 1);
    o.insert(o.end(), type, type + std::strlen(type) + 1);
    le32(o, (uint32_t)v.size());
    o.insert(o.end(), v.begin(), v.end());
}

bool write_exr(const std::string& path, const float* rgb, int w, int h) {
    std::vector<uint8_t> o;
    le32(o, 20000630u); // magic
    le32(o, 2u);        // version 2, single-part scanline
    std::vector<uint8_t> ch;
    for (const char* n : {"B", "G", "R"}) { // channels sorted by name
        ch.insert(ch.end(), n, n + 2);
        le32(ch, 2u); // FLOAT
        ch.insert(ch.end(), {0, 0, 0, 0});
        le32(ch, 1u);
        le32(ch, 1u);
    }
    ch.push_back(0);
    attr(o, "channels", "chlist", ch);
    attr(o, "compression", "compression", {0});
    std::vector<uint8_t> box;
    le32(box, 0u);
    le32(box, 0u);
    le32(box, (uint32_t)(w - 1));
    le32(box, (uint32_t)(h - 1));
    attr(o, "dataWindow", "box2i", box);
    attr(o, "displayWindow", "box2i", box);
    attr(o, "lineOrder", "lineOrder", {0});
    std::vector<uint8_t> par;
    const float one = 1.0f;
    uint32_t ob;
    std::memcpy(&ob, &one, 4);
    le32(par, ob);
    attr(o, "pixelAspectRatio", "float", par);
    std::vector<uint8_t> cen(8, 0);
    attr(o, "screenWindowCenter", "v2f", cen);
    attr(o, "screenWindowWidth", "float", par);
    o.push_back(0); // end of header
    const size_t line = 8 + (size_t)w * 3 * 4;
    const size_t table = o.size();
    for (int j = 0; j < h; ++j) le64(o, (uint64_t)(table + 8 * (size_t)h + line * j));
    for (int j = 0; j < h; ++j) {
        le32(o, (uint32_t)j);
        le32(o, (uint32_t)(w * 3 * 4));
        for (int c : {2, 1, 0}) // B, G, R planes of the scanline
            for (int i = 0; i < w; ++i) {
                uint32_t u;
                std::memcpy(&u, rgb + 3 * ((size_t)i + (size_t)w * j) + c, 4);
                le32(o, u);
            }
    }
    return write_file(path, o);
}

} // namespace

extern "C" int blz_write_image(const char* path, const float* rgb, int width, int height) {
    if (!path || !rgb || width <= 0 || height <= 0) {
        blz::set_error("blz_write_image: bad argument");
        return BLZ_E_INVALID;
    }
    const std::string p(path);
    const std::string ext = p.size() >= 4 ? p.substr(p.size() - 3) : "";
    bool ok = false;
    if (ext == "ppm") ok = write_ppm(p, rgb, width, height);
    else if (ext == "png") ok = write_png(p, rgb, width, height);
    else if (ext == "exr") ok = write_exr(p, rgb, width, height);
    else {
        blz::set_error("blz_write_image: extension must be ppm, png or exr: " + p);
        return BLZ_E_INVALID;
    }
    if (!ok) {
        blz::set_error("blz_write_image: cannot write " + p);
        return BLZ_E_IO;
    }
    return BLZ_OK;
}
