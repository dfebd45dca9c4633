// rt_image.cpp -- writes RGBA8 images from host memory: PNG (RFC 2083;
// zlib stream of stored deflate blocks, RFC 1950/1951, so no compressor is
// needed) or binary PPM (P6, alpha dropped).  The reference presents its
// target through the swapchain instead (src/main.cpp:155-161); the writer is
// how a headless frame loop shows its display image.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rt_common.hpp"

namespace {

uint32_t crc_table_entry(uint32_t c) {
    for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
    return c;
}

struct Crc32 {
    uint32_t table[256];
    Crc32() {
        for (uint32_t i = 0; i < 256; ++i) table[i] = crc_table_entry(i);
    }
    uint32_t update(uint32_t crc, const uint8_t* p, size_t n) const {
        crc = ~crc;
        for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
        return ~crc;
    }
};

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void put_chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
    static const Crc32 crc;
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc.update(0u, out.data() + start, out.size() - start));
}

bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suffix[i]) return false;
    return true;
}

int write_file(const char* path, const std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return RT_ERR_INVALID;
    const size_t wrote = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int closed = std::fclose(f);
    return wrote == bytes.size() && closed == 0 ? RT_OK : RT_ERR_INVALID;
}

std::vector<uint8_t> encode_png(const uint8_t* rgba, uint32_t w, uint32_t h) {
    // raw scanlines: filter type 0 then the row
    const size_t row = (size_t)w * 4;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + y * row, rgba + (y + 1) * row);
    }
    // zlib header (deflate, 32K window, no dictionary; FCHECK makes it % 31 == 0)
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);                // BFINAL, BTYPE 00 (stored)
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;                         // Adler-32 of the raw data
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);

    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});     // 8 bits, RGBA, deflate, filter 0, no interlace
    put_chunk(out, "IHDR", ihdr);
    put_chunk(out, "IDAT", z);
    put_chunk(out, "IEND", {});
    return out;
}

std::vector<uint8_t> encode_ppm(const uint8_t* rgba, uint32_t w, uint32_t h) {
    const std::string head = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    std::vector<uint8_t> out(head.begin(), head.end());
    out.reserve(out.size() + (size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; ++i) out.insert(out.end(), rgba + 4 * i, rgba + 4 * i + 3);
    return out;
}

}  // namespace

extern "C" int rt_image_write(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h) {
    if (!path || !rgba || w == 0 || h == 0) return RT_ERR_INVALID;
    if ((uint64_t)w * h > (1ull << 31)) return RT_ERR_INVALID;
    try {
        const std::string p(path);
        if (ends_with(p, ".png")) return write_file(path, encode_png(rgba, w, h));
        if (ends_with(p, ".ppm")) return write_file(path, encode_ppm(rgba, w, h));
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
    return RT_ERR_UNSUPPORTED;
}
