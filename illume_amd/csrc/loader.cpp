/*
 * loader.cpp - reference-schema scene JSON and OBJ meshes to an ilm_scene_desc (host only).
 *
 * Reads the schema of HostScene (src/scene/scene.cpp:11-99) and Renderer's render_settings
 * (src/renderer.cu:39-59): environment_map, camera, mediums, materials, meshes, primitives.
 * Field defaults follow the reference loaders: texture_from_json (src/texture.cu:7-71),
 * material_from_json (src/material.cu:7-97), medium_from_json (src/medium.cu:5-14),
 * camera_from_json (src/camera.cu:20-30), transform_from_json (src/math/transform.cu:28-44),
 * mesh_from_json (src/primitives/mesh.cu:430-443), Tonemapper (src/tonemapper.cu:5-30).
 * Fallbacks keep the reference's: unknown material type -> black diffuse, unknown conductor ->
 * "Cu", unknown mesh name -> mesh 0, missing resolution -> 512x512.
 * A camera without "depth_of_field" focuses at distance 1 (camera_from_json leaves the field
 * uninitialised; with aperture 0 the focus distance does not change ray directions, while 0
 * would make every camera ray direction 0/0).
 * Additive keys (no reference counterpart): material "eta"/"k" (explicit conductor IOR) and
 * "distribution": "ggx"; bitmap textures also read PFM (float) files. The reference reads images
 * through stb_image: here PNG, binary PPM/PGM (8-bit texels) and Radiance HDR (float texels)
 * are read by the decoders below, with stb's conventions for each.
 * The OBJ reader accepts faces with more than three vertices (fan triangulation) and long
 * lines; vertex merging assigns indices in first-occurrence order as src/primitives/mesh.cu:264-340.
 */
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include <zlib.h>

#include "../../include/illume_c.h"

namespace {

/* ---------------- minimal JSON ---------------- */
struct JV {
    enum T { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<JV> a;
    std::vector<std::pair<std::string, JV>> o;
    const JV* get(const char* k) const {
        if (t != OBJ) return nullptr;
        for (auto& kv : o) if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const char* p;
    const char* e;
    std::string err;
    void ws() {
        for (;;) {
            while (p < e && std::isspace((unsigned char)*p)) p++;
            if (p + 1 < e && p[0] == '/' && p[1] == '/') { while (p < e && *p != '\n') p++; continue; }
            break;
        }
    }
    bool parse(JV& v) {
        ws();
        if (p >= e) { err = "unexpected end"; return false; }
        char c = *p;
        if (c == '{') {
            v.t = JV::OBJ;
            p++;
            ws();
            if (p < e && *p == '}') { p++; return true; }
            for (;;) {
                ws();
                JV k;
                if (p >= e || *p != '"' || !str(k.s)) { err = "expected key"; return false; }
                ws();
                if (p >= e || *p != ':') { err = "expected ':'"; return false; }
                p++;
                JV val;
                if (!parse(val)) return false;
                v.o.emplace_back(k.s, std::move(val));
                ws();
                if (p < e && *p == ',') { p++; ws(); if (p < e && *p == '}') { p++; return true; } continue; }
                if (p < e && *p == '}') { p++; return true; }
                err = "expected ',' or '}'";
                return false;
            }
        }
        if (c == '[') {
            v.t = JV::ARR;
            p++;
            ws();
            if (p < e && *p == ']') { p++; return true; }
            for (;;) {
                JV val;
                if (!parse(val)) return false;
                v.a.push_back(std::move(val));
                ws();
                if (p < e && *p == ',') { p++; ws(); if (p < e && *p == ']') { p++; return true; } continue; }
                if (p < e && *p == ']') { p++; return true; }
                err = "expected ',' or ']'";
                return false;
            }
        }
        if (c == '"') { v.t = JV::STR; return str(v.s); }
        if (!strncmp(p, "true", 4)) { v.t = JV::BOOL; v.b = true; p += 4; return true; }
        if (!strncmp(p, "false", 5)) { v.t = JV::BOOL; v.b = false; p += 5; return true; }
        if (!strncmp(p, "null", 4)) { v.t = JV::NUL; p += 4; return true; }
        char* end = nullptr;
        v.n = std::strtod(p, &end);
        if (end == p) { err = "bad value"; return false; }
        v.t = JV::NUM;
        p = end;
        return true;
    }
    bool str(std::string& out) {
        p++;
        while (p < e && *p != '"') {
            if (*p == '\\' && p + 1 < e) { p++; char c = *p; out.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c); p++; continue; }
            out.push_back(*p++);
        }
        if (p >= e) { err = "unterminated string"; return false; }
        p++;
        return true;
    }
};

/* ---------------- owned descriptor ---------------- */
struct TexOwn {
    std::vector<float> f;
    std::vector<uint8_t> u8;
};

}  // namespace

struct ilm_desc {
    ilm_scene_desc desc;
    ilm_render_params settings;
    std::vector<ilm_mesh_desc> meshes;
    std::vector<std::vector<float>> pos, nrm, uv;
    std::vector<std::vector<int32_t>> idx;
    std::vector<ilm_instance_desc> instances;
    std::vector<ilm_sphere_desc> spheres;
    std::vector<ilm_material_desc> materials;
    std::vector<ilm_medium_desc> mediums;
    std::vector<std::unique_ptr<TexOwn>> tex;
    std::string err;
    double obj_parse_s = 0.0, obj_merge_s = 0.0;   /* wall time of the OBJ reads, summed over meshes */
};

static thread_local std::string g_loader_err;

namespace {

bool get_vec3(const JV* v, float out[3]) {
    if (!v) return false;
    if (v->t == JV::NUM) { out[0] = out[1] = out[2] = (float)v->n; return true; }
    if (v->t == JV::ARR && v->a.size() >= 3) {
        for (int i = 0; i < 3; i++) out[i] = (float)v->a[i].n;
        return true;
    }
    return false;
}
float get_num(const JV& o, const char* k, float def) {
    const JV* v = o.get(k);
    return (v && v->t == JV::NUM) ? (float)v->n : def;
}
bool get_bool(const JV& o, const char* k, bool def) {
    const JV* v = o.get(k);
    return (v && v->t == JV::BOOL) ? v->b : def;
}
std::string get_str(const JV& o, const char* k, const std::string& def = "") {
    const JV* v = o.get(k);
    return (v && v->t == JV::STR) ? v->s : def;
}

std::string dir_of(const std::string& path) {
    size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string(".") : path.substr(0, s);
}
std::string resolve(const std::string& file, const std::string& base) {
    std::ifstream f(file);
    if (f.good()) return file;
    std::string alt = base + "/" + file;
    std::ifstream g(alt);
    if (g.good()) return alt;
    /* reference scenes name paths from the repository root (e.g. "res/cube.obj") */
    std::string up = dir_of(dir_of(base)) + "/" + file;
    std::ifstream h(up);
    if (h.good()) return up;
    return file;
}

bool read_pfm(const std::string& path, int& w, int& h, std::vector<float>& rgba) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[3] = {0};
    double scale = 0;
    if (fscanf(f, "%2s %d %d %lf", magic, &w, &h, &scale) != 4 || (strcmp(magic, "PF") && strcmp(magic, "Pf")) || w <= 0 || h <= 0) { fclose(f); return false; }
    fgetc(f);
    int ch = magic[1] == 'F' ? 3 : 1;
    std::vector<float> raw((size_t)w * h * ch);
    if (fread(raw.data(), 4, raw.size(), f) != raw.size()) { fclose(f); return false; }
    fclose(f);
    bool little = scale < 0;
    rgba.assign((size_t)w * h * 4, 1.0f);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            /* PFM rows run bottom to top */
            const float* src = &raw[((size_t)(h - 1 - y) * w + x) * ch];
            for (int c = 0; c < 3; c++) {
                float v = src[ch == 3 ? c : 0];
                if (!little) { uint32_t u; memcpy(&u, &v, 4); u = __builtin_bswap32(u); memcpy(&v, &u, 4); }
                rgba[((size_t)y * w + x) * 4 + c] = v;
            }
        }
    return true;
}
bool read_file(const std::string& path, std::vector<uint8_t>& buf) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    buf.clear();
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    return true;
}

/* Largest bitmap texture read from a file, in texels. */
const int64_t MAX_TEXELS = (int64_t)1 << 28;

/* Binary PNM: P6 (RGB) and P5 (grey), maxval 1..255, '#' comments in the header. Bytes are
 * taken as they are (not rescaled by maxval), as stb_image's PNM reader gives them to
 * ImageIO::load_ldr (src/imageio.cpp:6-13). */
bool read_pnm(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgba, std::string& why) {
    size_t p = 2;
    auto field = [&](long& v) {
        for (;;) {
            while (p < b.size() && isspace(b[p])) p++;
            if (p < b.size() && b[p] == '#') { while (p < b.size() && b[p] != '\n') p++; continue; }
            break;
        }
        if (p >= b.size() || !isdigit(b[p])) return false;
        v = 0;
        while (p < b.size() && isdigit(b[p]) && v < (1L << 30)) v = v * 10 + (b[p++] - '0');
        return true;
    };
    if (b.size() < 2 || b[0] != 'P' || (b[1] != '6' && b[1] != '5')) { why = "not a binary PNM"; return false; }
    const int ch = b[1] == '6' ? 3 : 1;
    long W = 0, H = 0, mx = 0;
    if (!field(W) || !field(H) || !field(mx)) { why = "bad PNM header"; return false; }
    if (W <= 0 || H <= 0 || W * H > MAX_TEXELS) { why = "bad PNM size"; return false; }
    if (mx < 1 || mx > 255) { why = "PNM maxval above 255 is not read"; return false; }
    p++;   /* one whitespace byte ends the header */
    const size_t need = (size_t)W * H * ch;
    if (p + need > b.size()) { why = "PNM data truncated"; return false; }
    w = (int)W; h = (int)H;
    rgba.assign((size_t)W * H * 4, 255);
    for (size_t i = 0; i < (size_t)W * H; i++)
        for (int c = 0; c < 3; c++) rgba[4 * i + c] = b[p + i * ch + (ch == 3 ? c : 0)];
    return true;
}

/* Radiance RGBE (.hdr), the format ImageIO::load_hdr reads through stb_image's loader
 * (src/imageio.cpp:15-21; src/texture.cu:49-55 sends every path ending in "hdr" there):
 * a "#?RADIANCE" or "#?RGBE" first line, header lines up to an empty line, one of which must be
 * FORMAT=32-bit_rle_rgbe, then the resolution line "-Y <h> +X <w>" (rows top to bottom). Pixel
 * data is flat RGBE when w < 8 or w >= 32768, or when the first scanline does not start with
 * the adaptive run-length header (2, 2, w >> 8, w & 255); otherwise every scanline carries that
 * header and then its four channels one after the other, each as runs: a count byte above 128
 * repeats the next byte count - 128 times, a count of 1..128 is followed by that many literal
 * bytes. A texel is (r, g, b) * 2^(e - 136), all 0 when e = 0, alpha 1. A scanline after the
 * first without its run-length header is an error here (stb restarts the image as flat data). */
bool read_hdr(const std::vector<uint8_t>& b, int& w, int& h, std::vector<float>& rgba, std::string& why) {
    size_t p = 0;
    auto line = [&](std::string& out) {
        out.clear();
        while (p < b.size() && b[p] != '\n') out.push_back((char)b[p++]);
        if (p >= b.size()) return false;
        p++;
        return true;
    };
    std::string ln;
    if (!line(ln) || (ln != "#?RADIANCE" && ln != "#?RGBE")) { why = "not a Radiance HDR file"; return false; }
    bool fmt = false;
    for (;;) {
        if (!line(ln)) { why = "HDR header not terminated"; return false; }
        if (ln.empty()) break;
        if (ln == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    if (!fmt) { why = "HDR format is not 32-bit_rle_rgbe"; return false; }
    long H = 0, W = 0;
    if (!line(ln) || sscanf(ln.c_str(), "-Y %ld +X %ld", &H, &W) != 2) { why = "HDR orientation other than -Y +X"; return false; }
    if (W <= 0 || H <= 0 || W * H > MAX_TEXELS) { why = "bad HDR size"; return false; }
    w = (int)W; h = (int)H;
    rgba.assign((size_t)W * H * 4, 0.0f);
    auto put = [&](size_t i, const uint8_t* e) {
        float* o = &rgba[4 * i];
        if (e[3]) {
            const float f = ldexpf(1.0f, (int)e[3] - 136);
            o[0] = (float)e[0] * f; o[1] = (float)e[1] * f; o[2] = (float)e[2] * f;
        }
        o[3] = 1.0f;
    };
    const size_t n = (size_t)W * H;
    bool flat = W < 8 || W >= 32768;
    if (!flat && (p + 4 > b.size() || b[p] != 2 || b[p + 1] != 2 || (b[p + 2] & 0x80))) flat = true;
    if (flat) {
        if (p + 4 * n > b.size()) { why = "HDR data truncated"; return false; }
        for (size_t i = 0; i < n; i++) put(i, &b[p + 4 * i]);
        return true;
    }
    std::vector<uint8_t> row((size_t)W * 4);
    for (long y = 0; y < H; y++) {
        if (p + 4 > b.size() || b[p] != 2 || b[p + 1] != 2 || (b[p + 2] & 0x80)) { why = "HDR scanline without run-length header"; return false; }
        if (((long)b[p + 2] << 8 | b[p + 3]) != W) { why = "HDR scanline length differs from the width"; return false; }
        p += 4;
        for (int c = 0; c < 4; c++) {
            long x = 0;
            while (x < W) {
                if (p >= b.size()) { why = "HDR data truncated"; return false; }
                int cnt = b[p++];
                if (cnt > 128) {
                    cnt -= 128;
                    if (cnt > W - x || p >= b.size()) { why = "bad HDR run"; return false; }
                    const uint8_t v = b[p++];
                    for (int k = 0; k < cnt; k++) row[4 * (x++) + c] = v;
                } else {
                    if (cnt == 0 || cnt > W - x || p + cnt > b.size()) { why = "bad HDR run"; return false; }
                    for (int k = 0; k < cnt; k++) row[4 * (x++) + c] = b[p++];
                }
            }
        }
        for (long x = 0; x < W; x++) put((size_t)y * W + x, &row[4 * x]);
    }
    return true;
}

/* PNG to 8-bit RGBA, as stb_image's stbi_load(..., 4) gives ImageIO::load_ldr
 * (src/imageio.cpp:6-13): every colour type (grey, RGB, palette, grey + alpha, RGBA), bit depths
 * 1-16, both interlace methods, the five scanline filters. Grey below 8 bits is scaled to
 * 0..255 (times 255, 85 or 17); 16-bit samples keep their high byte; grey is replicated into
 * R, G and B; alpha is 255 where the image has none (tRNS is not applied: rendering reads RGB
 * only). IDAT data is inflated with zlib; chunk CRCs are not checked, and an unknown critical
 * chunk is an error, as in stb. */
uint32_t be32(const uint8_t* q) { return (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3]; }

bool read_png(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgba, std::string& why) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (b.size() < 8 || memcmp(b.data(), sig, 8)) { why = "not a PNG file"; return false; }
    size_t p = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> pal, z;
    bool ihdr = false, iend = false;
    while (!iend) {
        if (p + 12 > b.size()) { why = "PNG truncated"; return false; }
        const uint32_t len = be32(&b[p]);
        const uint8_t* t = &b[p + 4];
        if (len > b.size() - p - 12) { why = "PNG chunk past the end of the file"; return false; }
        const uint8_t* d = &b[p + 8];
        const std::string type((const char*)t, 4);
        if (!ihdr && type != "IHDR") { why = "PNG does not start with IHDR"; return false; }
        if (type == "IHDR") {
            if (len != 13 || ihdr) { why = "bad IHDR"; return false; }
            W = be32(d); H = be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
            if (d[10] != 0 || d[11] != 0 || interlace > 1) { why = "unknown PNG compression, filter or interlace method"; return false; }
            ihdr = true;
        } else if (type == "PLTE") {
            if (len % 3 || len > 768) { why = "bad PLTE"; return false; }
            pal.assign(d, d + len);
        } else if (type == "IDAT") {
            z.insert(z.end(), d, d + len);
        } else if (type == "IEND") {
            iend = true;
        } else if (!(t[0] & 32)) {
            why = "unknown critical PNG chunk " + type;
            return false;
        }
        p += 12 + (size_t)len;
    }
    const bool okdepth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                         (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                         ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!okdepth) { why = "bad PNG colour type / bit depth"; return false; }
    if (W == 0 || H == 0 || (int64_t)W * H > MAX_TEXELS) { why = "bad PNG size"; return false; }
    if (ctype == 3 && pal.empty()) { why = "palette PNG without PLTE"; return false; }
    const int chans = ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 1;
    const int bpp_bits = chans * depth;
    const size_t fbpp = (size_t)std::max(1, bpp_bits / 8);   /* filter distance in bytes */
    /* Adam7 passes: start x, start y, step x, step y (one pass of the whole image otherwise) */
    static const int A7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = interlace ? 7 : 1;
    size_t raw_len = 0;
    for (int k = 0; k < npass; k++) {
        const uint32_t pw = interlace ? (W - A7[k][0] + A7[k][2] - 1) / A7[k][2] * (W > (uint32_t)A7[k][0]) : W;
        const uint32_t ph = interlace ? (H - A7[k][1] + A7[k][3] - 1) / A7[k][3] * (H > (uint32_t)A7[k][1]) : H;
        if (pw && ph) raw_len += (size_t)ph * (1 + ((size_t)pw * bpp_bits + 7) / 8);
    }
    std::vector<uint8_t> raw(raw_len);
    {
        z_stream zs;
        memset(&zs, 0, sizeof(zs));
        if (inflateInit(&zs) != Z_OK) { why = "zlib inflateInit failed"; return false; }
        zs.next_in = z.data();
        zs.avail_in = (uInt)z.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt)raw.size();
        const int zr = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR && zr != Z_OK) || got < raw.size()) { why = "PNG image data inflates short or corrupt"; return false; }
    }
    w = (int)W; h = (int)H;
    rgba.assign((size_t)W * H * 4, 255);
    static const int gscale[17] = {0, 255, 85, 0, 17, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1};
    size_t off = 0;
    std::vector<uint8_t> prev, cur;
    for (int k = 0; k < npass; k++) {
        const uint32_t x0 = interlace ? A7[k][0] : 0, y0 = interlace ? A7[k][1] : 0;
        const uint32_t sx = interlace ? A7[k][2] : 1, sy = interlace ? A7[k][3] : 1;
        const uint32_t pw = W > x0 ? (W - x0 + sx - 1) / sx : 0, ph = H > y0 ? (H - y0 + sy - 1) / sy : 0;
        if (!pw || !ph) continue;
        const size_t stride = ((size_t)pw * bpp_bits + 7) / 8;
        prev.assign(stride, 0);
        cur.resize(stride);
        for (uint32_t r = 0; r < ph; r++) {
            const int ft = raw[off];
            const uint8_t* src = &raw[off + 1];
            off += 1 + stride;
            for (size_t i = 0; i < stride; i++) {
                const int a = i >= fbpp ? cur[i - fbpp] : 0, up = prev[i], c = i >= fbpp ? prev[i - fbpp] : 0;
                int pred;
                switch (ft) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = up; break;
                case 3: pred = (a + up) >> 1; break;
                case 4: {
                    const int pp = a + up - c, pa = abs(pp - a), pb = abs(pp - up), pc = abs(pp - c);
                    pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                    break;
                }
                default: why = "bad PNG filter type"; return false;
                }
                cur[i] = (uint8_t)(src[i] + pred);
            }
            const uint32_t y = y0 + r * sy;
            for (uint32_t q = 0; q < pw; q++) {
                const uint32_t x = x0 + q * sx;
                uint8_t* o = &rgba[((size_t)y * W + x) * 4];
                int s[4];
                for (int c = 0; c < chans; c++) {
                    if (depth == 8) s[c] = cur[(size_t)q * chans + c];
                    else if (depth == 16) s[c] = cur[((size_t)q * chans + c) * 2];   /* high byte */
                    else {
                        const size_t bit = (size_t)q * depth;
                        s[c] = (cur[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1 << depth) - 1);
                    }
                }
                if (ctype == 3) {
                    if ((size_t)s[0] * 3 + 2 >= pal.size()) { why = "PNG palette index out of range"; return false; }
                    o[0] = pal[3 * s[0]]; o[1] = pal[3 * s[0] + 1]; o[2] = pal[3 * s[0] + 2];
                } else if (chans <= 2) {
                    const int g = depth < 8 ? s[0] * gscale[depth] : s[0];
                    o[0] = o[1] = o[2] = (uint8_t)g;
                    if (chans == 2) o[3] = (uint8_t)s[1];
                } else {
                    o[0] = (uint8_t)s[0]; o[1] = (uint8_t)s[1]; o[2] = (uint8_t)s[2];
                    if (chans == 4) o[3] = (uint8_t)s[3];
                }
            }
            prev.swap(cur);
            cur.resize(stride);
        }
    }
    return true;
}

ilm_texture_desc tex_const(float r, float g, float b) {
    ilm_texture_desc t;
    memset(&t, 0, sizeof(t));
    t.type = ILM_TEX_CONSTANT;
    t.color[0] = r; t.color[1] = g; t.color[2] = b;
    t.scale[0] = t.scale[1] = 1.0f;
    return t;
}

/* src/texture.cu:7-71 */
ilm_texture_desc texture_from_json(const JV& j, ilm_desc* D, const std::string& base) {
    if (j.t == JV::NUM) return tex_const((float)j.n, (float)j.n, (float)j.n);
    if (j.t == JV::ARR) {
        float c[3] = {0, 0, 0};
        get_vec3(&j, c);
        return tex_const(c[0], c[1], c[2]);
    }
    std::string type = get_str(j, "type");
    if (type == "constant") {
        float c[3] = {0, 0, 0};
        get_vec3(j.get("color"), c);
        return tex_const(c[0], c[1], c[2]);
    }
    if (type == "checkerboard") {
        ilm_texture_desc t = tex_const(0, 0, 0);
        t.type = ILM_TEX_CHECKERBOARD;
        get_vec3(j.get("on_color"), t.on);
        get_vec3(j.get("off_color"), t.off);
        const JV* sc = j.get("scale");
        if (sc && sc->t == JV::ARR && sc->a.size() >= 2) { t.scale[0] = (float)sc->a[0].n; t.scale[1] = (float)sc->a[1].n; }
        return t;
    }
    if (type == "bitmap") {
        std::string file = get_str(j, "file");
        if (file.empty()) return tex_const(0, 0, 0);
        std::string path = resolve(file, base);
        ilm_texture_desc t = tex_const(0, 0, 0);
        const JV* sc = j.get("scale");
        std::unique_ptr<TexOwn> own(new TexOwn());
        int w = 0, h = 0;
        bool ok = false;
        std::string why = "cannot open the file";
        std::vector<uint8_t> bytes;
        const bool opened = read_file(path, bytes);
        if (path.size() > 4 && path.substr(path.size() - 4) == ".pfm") {
            ok = read_pfm(path, w, h, own->f);
            if (!ok) why = "not a readable PFM file";
            t.is_hdr = 1;
        } else if (opened && path.size() >= 3 && path.substr(path.size() - 3) == "hdr") {
            /* float texels, bilinear, as src/texture.cu:49-55 */
            ok = read_hdr(bytes, w, h, own->f, why);
            t.is_hdr = 1;
        } else if (opened) {
            /* 8-bit texels, point sampled, as src/texture.cu:57-63; the format is told by content */
            if (bytes.size() >= 8 && bytes[0] == 0x89 && bytes[1] == 'P') ok = read_png(bytes, w, h, own->u8, why);
            else if (bytes.size() >= 2 && bytes[0] == 'P' && (bytes[1] == '6' || bytes[1] == '5')) ok = read_pnm(bytes, w, h, own->u8, why);
            else why = "unsupported image format (PNG, binary PPM/PGM, Radiance HDR and PFM are read)";
            t.is_hdr = 0;
        }
        if (!ok) {
            fprintf(stderr, "texture_from_json: cannot read %s (%s), using black\n", path.c_str(), why.c_str());
            return tex_const(0, 0, 0);
        }
        t.type = ILM_TEX_BITMAP;
        t.width = w;
        t.height = h;
        t.texels_f32 = own->f.empty() ? nullptr : own->f.data();
        t.texels_u8 = own->u8.empty() ? nullptr : own->u8.data();
        if (sc && sc->t == JV::ARR && sc->a.size() >= 2) { t.scale[0] = (float)sc->a[0].n; t.scale[1] = (float)sc->a[1].n; }
        D->tex.push_back(std::move(own));
        return t;
    }
    fprintf(stderr, "texture_from_json: invalid texture type %s\n", type.c_str());
    return tex_const(0, 0, 0);
}

ilm_transform_desc transform_from_json(const JV* j) {
    ilm_transform_desc t;
    memset(&t, 0, sizeof(t));
    t.scale[0] = t.scale[1] = t.scale[2] = 1.0f;
    if (!j) return t;
    get_vec3(j->get("translation"), t.translation);
    get_vec3(j->get("scale"), t.scale);
    get_vec3(j->get("rotation"), t.rotation);
    if (j->get("lookat")) { t.has_lookat = 1; get_vec3(j->get("lookat"), t.lookat); }
    return t;
}

/* src/conductor.h:18-25 (refractiveindex.info values at 700, 510, 440 nm) */
struct CondEntry { const char* name; float eta[3]; float k[3]; };
const CondEntry kConductors[] = {
    {"Ag", {0.16761f, 0.14032f, 0.13529f}, {4.2868f, 2.8794f, 2.2916f}},
    {"Al", {1.66130f, 0.82373f, 0.58461f}, {8.0444f, 5.9742f, 5.1875f}},
    {"Au", {0.14311f, 0.37495f, 1.44247f}, {3.9831f, 2.3857f, 1.6032f}},
    {"Cu", {0.21249f, 0.97909f, 1.3343f}, {4.1005f, 2.3691f, 2.2965f}},
    {"Pb", {1.78000f, 1.74000f, 1.4400f}, {3.5700f, 3.3100f, 3.1800f}},
};
bool conductor_get(const std::string& name, float eta[3], float k[3]) {
    for (const CondEntry& c : kConductors)
        if (name == c.name) { for (int i = 0; i < 3; i++) { eta[i] = c.eta[i]; k[i] = c.k[i]; } return true; }
    return false;
}

/* OBJ: src/primitives/mesh.cu:159-340, 363-388 */
/* OBJ face token "p", "p/t", "p//n" or "p/t/n" (1-based, negative = relative to the end);
 * false when it does not start with an index. */
bool obj_face_token(const char*& s, long& p, long& t, long& n) {
    char* e;
    p = std::strtol(s, &e, 10);
    if (e == s) return false;
    s = e;
    t = n = 0;
    if (*s == '/') {
        s++;
        if (*s != '/') { t = std::strtol(s, &e, 10); s = e; }
        if (*s == '/') { s++; n = std::strtol(s, &e, 10); s = e; }
    }
    while (*s && *s != ' ' && *s != '\t' && *s != '\n' && *s != '\r') s++;
    return true;
}

bool load_obj(const std::string& path, bool z_up, bool face_normals, bool has_tex, bool flip, ilm_desc* D, int mi) {
    const auto t_start = std::chrono::steady_clock::now();
    /* the whole file in memory, parsed in place (any line length) */
    std::string buf;
    {
        FILE* f = fopen(path.c_str(), "rb");
        if (!f) return false;
        if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return false; }
        const long size = ftell(f);
        if (size < 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return false; }
        buf.resize((size_t)size);
        const bool ok = fread(&buf[0], 1, (size_t)size, f) == (size_t)size;
        fclose(f);
        if (!ok) return false;
    }
    std::vector<float> v, vt, vn;
    struct FV { int p, t, n; };
    std::vector<FV> tris, poly;
    auto fix = [](long i, size_t n) -> int { return i > 0 ? (int)(i - 1) : (i < 0 ? (int)((long)n + i) : -1); };
    const char* s = buf.c_str();
    while (*s) {
        while (*s == ' ' || *s == '\t') s++;
        const char* line = s;
        while (*s && *s != '\n') s++;          /* s: end of this line */
        const char* q = line;
        char* e;
        if (q[0] == 'v' && (q[1] == ' ' || q[1] == '\t')) {
            float c[3];
            int k = 0;
            q += 2;
            for (; k < 3; k++) {
                c[k] = std::strtof(q, &e);
                if (e == q || e > s) break;
                q = e;
            }
            if (k == 3) {
                if (z_up) { v.push_back(c[0]); v.push_back(c[2]); v.push_back(c[1]); }
                else { v.push_back(c[0]); v.push_back(c[1]); v.push_back(c[2]); }
            }
        } else if (q[0] == 'v' && q[1] == 't') {
            float c[2] = {0, 0};
            q += 2;
            for (int k = 0; k < 2; k++) {
                const float x = std::strtof(q, &e);
                if (e == q || e > s) break;
                c[k] = x;
                q = e;
            }
            vt.push_back(c[0]); vt.push_back(c[1]);
        } else if (q[0] == 'v' && q[1] == 'n') {
            float c[3];
            int k = 0;
            q += 2;
            for (; k < 3; k++) {
                c[k] = std::strtof(q, &e);
                if (e == q || e > s) break;
                q = e;
            }
            if (k == 3) { vn.push_back(c[0]); vn.push_back(c[1]); vn.push_back(c[2]); }
        } else if (q[0] == 'f' && (q[1] == ' ' || q[1] == '\t')) {
            poly.clear();
            q += 2;
            for (;;) {
                while (q < s && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
                if (q >= s) break;
                long a, b, c;
                if (!obj_face_token(q, a, b, c)) {   /* not an index: skip the token */
                    while (q < s && *q != ' ' && *q != '\t') q++;
                    continue;
                }
                poly.push_back({fix(a, v.size() / 3), b ? fix(b, vt.size() / 2) : -1, c ? fix(c, vn.size() / 3) : -1});
            }
            for (size_t k = 2; k < poly.size(); k++) { tris.push_back(poly[0]); tris.push_back(poly[k - 1]); tris.push_back(poly[k]); }
        }
        if (*s == '\n') s++;
    }
    const auto t_parsed = std::chrono::steady_clock::now();
    D->obj_parse_s += std::chrono::duration<double>(t_parsed - t_start).count();
    /* merge identical (position, normal, uv) vertices in first-occurrence order */
    struct Key { uint32_t w[8]; bool operator==(const Key& o) const { return !memcmp(w, o.w, sizeof(w)); } };
    struct KH { size_t operator()(const Key& k) const { size_t h = 1469598103934665603ull; for (uint32_t x : k.w) h = (h ^ x) * 1099511628211ull; return h; } };
    std::unordered_map<Key, int, KH> map;
    map.reserve(tris.size() / 2 + 16);
    std::vector<float>& P = D->pos[mi];
    std::vector<float>& N = D->nrm[mi];
    std::vector<float>& T = D->uv[mi];
    std::vector<int32_t>& I = D->idx[mi];
    for (const FV& f : tris) {
        if (f.p < 0 || 3 * (size_t)f.p + 2 >= v.size() + 0) { if (f.p < 0 || (size_t)f.p >= v.size() / 3) return false; }
        float pos[3] = {v[3 * f.p], v[3 * f.p + 1], v[3 * f.p + 2]};
        float nrm[3] = {0, 0, 0};
        float tc[2] = {0, 0};
        if (!face_normals) {
            if (f.n < 0 || (size_t)f.n >= vn.size() / 3) return false;
            float n0 = vn[3 * f.n], n1 = vn[3 * f.n + 1], n2 = vn[3 * f.n + 2];
            float inv = 1.0f / std::sqrt((n0 * n0) + (n1 * n1) + (n2 * n2));
            nrm[0] = n0 * inv; nrm[1] = n1 * inv; nrm[2] = n2 * inv;
        }
        if (has_tex) {
            if (f.t < 0 || (size_t)f.t >= vt.size() / 2) return false;
            tc[0] = vt[2 * f.t]; tc[1] = vt[2 * f.t + 1];
        }
        Key k;
        memcpy(&k.w[0], pos, 12);
        memcpy(&k.w[3], nrm, 12);
        memcpy(&k.w[6], tc, 8);
        auto it = map.find(k);
        int id;
        if (it == map.end()) {
            id = (int)(P.size() / 3);
            map.emplace(k, id);
            P.insert(P.end(), pos, pos + 3);
            N.insert(N.end(), nrm, nrm + 3);
            T.insert(T.end(), tc, tc + 2);
        } else {
            id = it->second;
        }
        I.push_back(id);
    }
    if (flip) {
        if (face_normals) for (size_t t = 0; t + 2 < I.size(); t += 3) std::swap(I[t + 1], I[t + 2]);
        else for (float& x : N) x = -x;
    }
    D->obj_merge_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_parsed).count();
    return true;
}

}  // namespace

extern "C" {

int ilm_desc_load_times(const ilm_desc* d, double* parse_s, double* merge_s) {
    if (!d) return ILM_E_INVALID;
    if (parse_s) *parse_s = d->obj_parse_s;
    if (merge_s) *merge_s = d->obj_merge_s;
    return ILM_OK;
}

int ilm_desc_load_json(const char* path, ilm_desc** out) {
    if (!path || !out) return ILM_E_INVALID;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) { g_loader_err = std::string("cannot read ") + path; return ILM_E_IO; }
    std::stringstream buf;
    buf << f.rdbuf();
    std::string text = buf.str();
    Parser ps{text.data(), text.data() + text.size(), ""};
    JV root;
    if (!ps.parse(root) || root.t != JV::OBJ) {
        g_loader_err = "JSON parse error at offset " + std::to_string(ps.p - text.data()) + ": " + ps.err;
        return ILM_E_PARSE;
    }
    std::unique_ptr<ilm_desc> D(new ilm_desc());
    const std::string base = dir_of(path);
    memset(&D->desc, 0, sizeof(D->desc));
    memset(&D->settings, 0, sizeof(D->settings));

    if (const JV* env = root.get("environment_map")) {
        D->desc.has_environment = 1;
        D->desc.environment = texture_from_json(*env, D.get(), base);
    }
    ilm_camera_desc& cam = D->desc.camera;
    cam.transform = transform_from_json(nullptr);
    if (const JV* c = root.get("camera")) {
        cam.fov = get_num(*c, "fov", 0);
        cam.dof = get_num(*c, "depth_of_field", 1);
        cam.aperture = get_num(*c, "aperture_radius", 0);
        cam.transform = transform_from_json(c->get("transform"));
    }
    std::map<std::string, int> medium_idx;
    if (const JV* ms = root.get("mediums")) {
        for (const JV& m : ms->a) {
            ilm_medium_desc md;
            memset(&md, 0, sizeof(md));
            float ab[3] = {0, 0, 0};
            get_vec3(m.get("absorption"), ab);
            for (int i = 0; i < 3; i++) md.sigma_a[i] = 1.0f - ab[i];   /* src/medium.cu:9 */
            md.sigma_s = get_num(m, "scattering", 0);
            md.g = get_num(m, "g", 0);
            medium_idx[get_str(m, "name")] = (int)D->mediums.size();
            D->mediums.push_back(md);
        }
    }
    std::map<std::string, int> mat_idx;
    auto black_diffuse = [&]() {
        ilm_material_desc m;
        memset(&m, 0, sizeof(m));
        m.type = ILM_MAT_DIFFUSE;
        m.albedo = tex_const(0, 0, 0);
        m.roughness = tex_const(1e-4f, 1e-4f, 1e-4f);
        m.medium = -1;
        return m;
    };
    if (const JV* ms = root.get("materials")) {
        for (const JV& j : ms->a) {
            ilm_material_desc m = black_diffuse();
            std::string type = get_str(j, "type");
            if (const JV* a = j.get("albedo")) m.albedo = texture_from_json(*a, D.get(), base);
            if (const JV* r = j.get("roughness")) m.roughness = texture_from_json(*r, D.get(), base);
            auto mi = medium_idx.find(get_str(j, "medium"));
            int medium = mi == medium_idx.end() ? -1 : mi->second;
            m.ior = get_num(j, "ior", 0);
            m.distribution = get_str(j, "distribution") == "ggx" ? ILM_DIST_GGX : ILM_DIST_BECKMANN;
            if (type == "emissive") m.type = ILM_MAT_EMISSIVE;
            else if (type == "lambert") m.type = ILM_MAT_DIFFUSE;
            else if (type == "reflective") m.type = ILM_MAT_SPECULAR;
            else if (type == "refractive") { m.type = ILM_MAT_REFRACTIVE; m.medium = medium; }
            else if (type == "roughreflective") m.type = ILM_MAT_ROUGHREFLECTIVE;
            else if (type == "roughrefractive") { m.type = ILM_MAT_ROUGHREFRACTIVE; m.medium = medium; }
            else if (type == "conductor" || type == "roughconductor") {
                m.type = type == "conductor" ? ILM_MAT_CONDUCTOR : ILM_MAT_ROUGHCONDUCTOR;
                if (type == "conductor") m.roughness = tex_const(1e-4f, 1e-4f, 1e-4f);
                m.albedo = tex_const(0, 0, 0);
                if (j.get("eta") && j.get("k")) {
                    get_vec3(j.get("eta"), m.eta);
                    get_vec3(j.get("k"), m.k);
                } else if (!conductor_get(get_str(j, "material"), m.eta, m.k)) {
                    fprintf(stderr, "material_from_json: invalid conductor material, defaulting to copper\n");
                    conductor_get("Cu", m.eta, m.k);
                }
            } else if (type == "plastic") m.type = ILM_MAT_PLASTIC;
            else {
                fprintf(stderr, "material_from_json: invalid material type %s\n", type.c_str());
                m = black_diffuse();
            }
            /* material_base resets roughness for every type that does not take one (src/material.cu:99-109) */
            if (m.type != ILM_MAT_ROUGHREFLECTIVE && m.type != ILM_MAT_ROUGHREFRACTIVE && m.type != ILM_MAT_ROUGHCONDUCTOR)
                m.roughness = tex_const(1e-4f, 1e-4f, 1e-4f);
            mat_idx[get_str(j, "name")] = (int)D->materials.size();
            D->materials.push_back(m);
        }
    }
    std::map<std::string, int> mesh_idx;
    if (const JV* ms = root.get("meshes")) {
        for (const JV& j : ms->a) {
            int mi = (int)D->meshes.size();
            D->pos.emplace_back(); D->nrm.emplace_back(); D->uv.emplace_back(); D->idx.emplace_back();
            ilm_mesh_desc md;
            memset(&md, 0, sizeof(md));
            md.face_normals = get_bool(j, "face_normals", true);
            md.has_texcoords = get_bool(j, "has_texcoords", false);
            md.leaf_size = (int)get_num(j, "bvh_items_per_node", 4);
            std::string file = resolve(get_str(j, "file"), base);
            if (!load_obj(file, get_bool(j, "z_up", false), md.face_normals, md.has_texcoords, get_bool(j, "flip_normals", false), D.get(), mi)) {
                g_loader_err = "cannot load mesh " + file;
                return ILM_E_IO;
            }
            if (!mesh_idx.count(get_str(j, "name"))) mesh_idx[get_str(j, "name")] = mi;
            D->meshes.push_back(md);
        }
    }
    if (const JV* ps2 = root.get("primitives")) {
        for (const JV& j : ps2->a) {
            std::string type = get_str(j, "type");
            std::string mname = get_str(j, "material");
            int mat;
            auto it = mat_idx.find(mname);
            if (it != mat_idx.end()) mat = it->second;
            else {
                mat = (int)D->materials.size();
                D->materials.push_back(black_diffuse());
                mat_idx[mname] = mat;
            }
            if (type == "sphere") {
                ilm_sphere_desc s;
                memset(&s, 0, sizeof(s));
                s.radius = get_num(j, "radius", 0);
                get_vec3(j.get("center"), s.center);
                s.material = mat;
                D->spheres.push_back(s);
            } else if (type == "mesh_instance") {
                ilm_instance_desc in;
                memset(&in, 0, sizeof(in));
                auto mit = mesh_idx.find(get_str(j, "mesh"));
                in.mesh = mit == mesh_idx.end() ? 0 : mit->second;
                in.material = mat;
                in.transform = transform_from_json(j.get("transform"));
                if (D->meshes.empty()) { g_loader_err = "mesh_instance without meshes"; return ILM_E_PARSE; }
                D->instances.push_back(in);
            }
        }
    }
    /* render_settings (src/renderer.cu:39-59, src/tonemapper.cu:5-30) */
    ilm_render_params& rp = D->settings;
    rp.width = rp.height = 512;
    rp.spp = 1;
    rp.max_depth = 1;
    rp.ray_bias = 0.0f;
    rp.tonemap = ILM_TONEMAP_LINEAR;
    if (const JV* rs = root.get("render_settings")) {
        rp.ray_bias = get_num(*rs, "ray_bias", 0.0f);
        const JV* res = rs->get("resolution");
        if (res && res->t == JV::ARR && res->a.size() >= 2) { rp.width = (int)res->a[0].n; rp.height = (int)res->a[1].n; }
        std::string op = get_str(*rs, "tonemapper", "linear");
        rp.tonemap = op == "Reinhard" ? ILM_TONEMAP_REINHARD : op == "Filmic" ? ILM_TONEMAP_FILMIC
                   : op == "Uncharted" ? ILM_TONEMAP_UNCHARTED : ILM_TONEMAP_LINEAR;
    }
    for (size_t i = 0; i < D->meshes.size(); i++) {
        ilm_mesh_desc& md = D->meshes[i];
        md.num_vertices = (int32_t)(D->pos[i].size() / 3);
        md.num_triangles = (int32_t)(D->idx[i].size() / 3);
        md.positions = D->pos[i].data();
        md.normals = md.face_normals ? nullptr : D->nrm[i].data();
        md.texcoords = md.has_texcoords ? D->uv[i].data() : nullptr;
        md.indices = D->idx[i].data();
    }
    ilm_scene_desc& sd = D->desc;
    sd.num_meshes = (int32_t)D->meshes.size();
    sd.meshes = D->meshes.data();
    sd.num_instances = (int32_t)D->instances.size();
    sd.instances = D->instances.data();
    sd.num_spheres = (int32_t)D->spheres.size();
    sd.spheres = D->spheres.data();
    sd.num_materials = (int32_t)D->materials.size();
    sd.materials = D->materials.data();
    sd.num_mediums = (int32_t)D->mediums.size();
    sd.mediums = D->mediums.data();
    *out = D.release();
    return ILM_OK;
}

const ilm_scene_desc* ilm_desc_scene(const ilm_desc* d) { return d ? &d->desc : nullptr; }
int ilm_desc_settings(const ilm_desc* d, ilm_render_params* out) {
    if (!d || !out) return ILM_E_INVALID;
    *out = d->settings;
    return ILM_OK;
}
void ilm_desc_free(ilm_desc* d) { delete d; }
const char* ilm_loader_error(void) { return g_loader_err.c_str(); }

}  /* extern "C" */
