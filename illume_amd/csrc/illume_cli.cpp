/*
 * illume_cli.cpp - command line of the reference (src/main.cu:12-43):
 *     illume <save> <scene.json> <spp> <maxdepth>
 * renders the scene and writes <save>-<w>x<h>-<spp>spp-<maxdepth>md.png, and beside it the
 * same name with .pfm: the mean radiance per pixel (sum of the samples / spp) as float RGB,
 * before tonemapping. The PNG writer is self-contained (zlib stored blocks, CRC-32 and Adler-32
 * computed here), replacing the reference's libpng output (src/bitmap.c:67-151); 8-bit RGB as
 * there. PFM: "PF" header, little-endian (negative scale), rows bottom to top.
 */
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "renderer.hpp"

static uint32_t crc_table[256];
static void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}
static uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}
static void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(x >> 24); v.push_back(x >> 16); v.push_back(x >> 8); v.push_back(x);
}
static void chunk(FILE* f, const char* type, const std::vector<uint8_t>& data) {
    std::vector<uint8_t> buf;
    put32(buf, (uint32_t)data.size());
    buf.insert(buf.end(), type, type + 4);
    buf.insert(buf.end(), data.begin(), data.end());
    uint32_t c = crc32(buf.data() + 4, buf.size() - 4) ^ 0xffffffffu;
    put32(buf, c);
    fwrite(buf.data(), 1, buf.size(), f);
}

/* 8-bit RGB PNG with an uncompressed (stored) zlib stream */
bool write_png(const char* path, const illume::Bitmap* b) {
    crc_init();
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    fwrite(sig, 1, 8, f);
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)b->width);
    put32(ihdr, (uint32_t)b->height);
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    chunk(f, "IHDR", ihdr);
    std::vector<uint8_t> raw;
    raw.reserve((size_t)b->height * (1 + 3 * (size_t)b->width));
    for (int y = 0; y < b->height; y++) {
        raw.push_back(0);
        for (int x = 0; x < b->width; x++) {
            const illume::Pixel& p = b->pixels[(size_t)y * b->width + x];
            raw.push_back((uint8_t)p.red); raw.push_back((uint8_t)p.green); raw.push_back((uint8_t)p.blue);
        }
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, bsum = 0;
    for (uint8_t v : raw) { a = (a + v) % 65521; bsum = (bsum + a) % 65521; }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back(n & 0xff); z.push_back(n >> 8);
        z.push_back(~n & 0xff); z.push_back((~n >> 8) & 0xff);
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (raw.empty()) break;
    }
    put32(z, (bsum << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    return fclose(f) == 0;
}

/* float RGB PFM of the mean radiance (rgb_sum / spp), little-endian, rows bottom to top */
bool write_pfm(const char* path, int w, int h, const std::vector<float>& rgb_sum, int spp) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
    std::vector<float> row(3 * (size_t)w);
    bool ok = true;
    for (int y = h - 1; y >= 0 && ok; y--) {
        for (size_t i = 0; i < 3 * (size_t)w; i++) row[i] = rgb_sum[3 * (size_t)y * w + i] / (float)spp;
        ok = fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
    }
    return (fclose(f) == 0) && ok;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        printf("illume: usage - <save> <scene> <spp> <maxdepth>\n");
        return 0;
    }
    int spp = (int)strtol(argv[3], nullptr, 10);
    int max_depth = (int)strtol(argv[4], nullptr, 10);
    try {
        illume::Device dev(0);
        illume::HostScene scene(dev, argv[2]);
        illume::Renderer renderer(scene, spp, max_depth);
        illume::Bitmap* image = illume::bitmap_new(renderer.get_width(), renderer.get_height());
        std::vector<float> radiance(3 * (size_t)renderer.get_width() * renderer.get_height());
        renderer.render_to_bitmap(image, radiance.data());
        printf("Rays per sec: %f rays/s\n", renderer.rays_per_sec());
        printf("Render time: %f seconds\n", renderer.last_info().device_ms * 1e-3);
        std::string name = std::string(argv[1]) + "-" + std::to_string(renderer.get_width()) + "x" +
                           std::to_string(renderer.get_height()) + "-" + std::to_string(spp) + "spp-" +
                           std::to_string(max_depth) + "md.png";
        if (!write_png(name.c_str(), image)) {
            fprintf(stderr, "illume: cannot write %s\n", name.c_str());
            illume::bitmap_free(image);
            return 1;
        }
        printf("Saved to: %s\n", name.c_str());
        const std::string pfm = name.substr(0, name.size() - 4) + ".pfm";
        if (!write_pfm(pfm.c_str(), renderer.get_width(), renderer.get_height(), radiance, renderer.get_spp())) {
            fprintf(stderr, "illume: cannot write %s\n", pfm.c_str());
            illume::bitmap_free(image);
            return 1;
        }
        printf("Saved to: %s\n", pfm.c_str());
        illume::bitmap_free(image);
    } catch (const illume::Error& e) {
        fprintf(stderr, "illume: %s (%d)\n", e.what(), e.code);
        return 1;
    }
    return 0;
}
