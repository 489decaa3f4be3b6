/*
 * renderer.hpp - the reference's C++ render interface over the C ABI (header only).
 *
 * Mirrors, for a caller of the reference:
 *   class Renderer { Renderer(json, HostScene&, int spp, int max_depth);
 *                    void render_to_bitmap(Bitmap*); int get_width(); int get_height(); }
 *                                                   src/renderer.h:9-26, src/renderer.cu:39-59,498-597
 *   HostScene(rapidjson::Document&)                 src/scene/scene.cpp:11-99
 *   Bitmap / Pixel {size_t red, green, blue}        src/bitmap.h:6-25
 * Differences a caller sees: the scene is read from a JSON path (the loader is the product's
 * own reader of the same schema), the device scene is uploaded once per HostScene and stays
 * resident across renders (the reference re-uploads it in every render_to_bitmap), and errors
 * throw illume::Error instead of exit(EXIT_FAILURE) (src/error_check.h:4-13).
 */
#pragma once
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/illume_c.h"

namespace illume {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, ilm_ctx* ctx = nullptr) {
    if (rc != ILM_OK) throw Error(rc, ctx ? ilm_last_error(ctx) : ilm_global_error());
}

/* Pixel {size_t red, green, blue} and Bitmap {width, height, pixels} of src/bitmap.h */
struct Pixel {
    size_t red, green, blue;
};
static_assert(sizeof(Pixel) == sizeof(ilm_pixel), "Pixel and ilm_pixel share one layout");

struct Bitmap {
    int width, height;
    Pixel* pixels;
};
inline Bitmap* bitmap_new(int w, int h) {
    Bitmap* b = new Bitmap{w, h, new Pixel[(size_t)w * h]()};
    return b;
}
inline void bitmap_free(Bitmap* b) {
    if (b) { delete[] b->pixels; delete b; }
}

/* One GPU context shared by the scenes and renderers of a process. */
class Device {
  public:
    explicit Device(int device = 0, const ilm_opts* opts = nullptr) { check(ilm_open(device, opts, &ctx_)); }
    ~Device() { ilm_close(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    ilm_ctx* get() const { return ctx_; }

  private:
    ilm_ctx* ctx_ = nullptr;
};

class HostScene {
  public:
    HostScene(Device& dev, const std::string& json_path) : dev_(dev) {
        int rc = ilm_desc_load_json(json_path.c_str(), &desc_);
        if (rc) throw Error(rc, ilm_loader_error());
        check(ilm_scene_upload(dev_.get(), ilm_desc_scene(desc_), &scene_), dev_.get());
    }
    ~HostScene() {
        ilm_scene_free(scene_);
        ilm_desc_free(desc_);
    }
    HostScene(const HostScene&) = delete;
    HostScene& operator=(const HostScene&) = delete;
    ilm_scene* device_scene() const { return scene_; }
    const ilm_desc* desc() const { return desc_; }
    Device& device() const { return dev_; }

  private:
    Device& dev_;
    ilm_desc* desc_ = nullptr;
    ilm_scene* scene_ = nullptr;
};

class Renderer {
  public:
    /* render_settings (resolution, ray_bias, tonemapper) come from the scene JSON, as in
     * src/renderer.cu:39-59 */
    Renderer(HostScene& scene, int spp, int max_depth) : scene_(scene) {
        check(ilm_desc_settings(scene.desc(), &params_));
        params_.spp = spp;
        params_.max_depth = max_depth;
        params_.first_sample = 0;
    }
    void render_to_bitmap(Bitmap* bitmap) { render_to_bitmap(bitmap, nullptr); }
    /* Also returns the Kahan-summed radiance of the spp samples, 3 floats per pixel (nullable). */
    void render_to_bitmap(Bitmap* bitmap, float* accum_rgb) {
        if (!bitmap || bitmap->width != params_.width || bitmap->height != params_.height)
            throw Error(ILM_E_INVALID, "bitmap size differs from the scene resolution");
        check(ilm_render(scene_.device().get(), scene_.device_scene(), &params_, accum_rgb,
                         reinterpret_cast<ilm_pixel*>(bitmap->pixels), &info_),
              scene_.device().get());
    }
    int get_spp() const { return params_.spp; }
    int get_width() const { return params_.width; }
    int get_height() const { return params_.height; }
    const ilm_render_info& last_info() const { return info_; }
    /* "Rays per sec" of src/renderer.cu:568-572, over the device time of the render loop */
    double rays_per_sec() const { return info_.device_ms > 0 ? info_.segments / (info_.device_ms * 1e-3) : 0.0; }

  private:
    HostScene& scene_;
    ilm_render_params params_;
    ilm_render_info info_{};
};

}  // namespace illume
