// omr_io.cpp -- C ABI of the file-format readers and writers (host only).
// The parsing itself is the drop-in's (dropin/Loaders.cpp): Model,
// HDRImage, SceneFromExternalFile, which mirror the reference's classes.
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/omr_capi.h"
#include "../dropin/ImageCodecs.h"
#include "../dropin/OmochiRenderer.h"

namespace {
thread_local std::string g_io_err;
// ImageHandler's table is process-wide; the loaders that fill it take turns
std::mutex g_images_mu;
int io_fail(int code, const std::string &msg) {
  g_io_err = msg;
  return code;
}

// A scene read from a file and the upload arrays flattened from it.
struct Loaded {
  std::unique_ptr<OmochiRenderer::Scene> scene;
  OmochiRenderer::SceneArrays arrays;
};

// A scene holding one placed OBJ model.
class ObjScene : public OmochiRenderer::Scene {
 public:
  bool Load(const std::string &path, const OmochiRenderer::Vector3 &pos, const OmochiRenderer::Vector3 &scale,
            const OmochiRenderer::Vector3 &rot, bool flip_v) {
    using OmochiRenderer::Matrix;
    using OmochiRenderer::Vector3;
    std::unique_ptr<OmochiRenderer::Model> m(new OmochiRenderer::Model);
    if (!m->ReadFromObj(path, flip_v)) return false;
    m->Transform(pos, scale,
                 Matrix::RotateAroundVector(Vector3(0, 0, 1), rot.z) * Matrix::RotateAroundVector(Vector3(0, 1, 0), rot.y) *
                     Matrix::RotateAroundVector(Vector3(1, 0, 0), rot.x));
    AddModel(m.release(), true, true);
    return true;
  }
};

int finish(Loaded *L, omr_loaded_scene *out) {
  L->scene->Describe(L->arrays);
  out->desc = L->arrays.desc;
  out->n_objects = L->arrays.prims.size();
  out->internal = L;
  return OMR_OK;
}
}  // namespace

extern "C" const char *omr_io_last_error(void) { return g_io_err.c_str(); }

extern "C" void omr_free(void *p) { std::free(p); }

extern "C" int omr_hdr_read(const char *path, uint32_t *width, uint32_t *height, float **rgb) {
  if (!path || !width || !height || !rgb) return io_fail(OMR_ERR_INVALID, "null argument");
  *rgb = nullptr;
  OmochiRenderer::HDRImage img;
  if (!img.ReadFromRadianceFile(path)) return io_fail(OMR_ERR_INVALID, std::string("cannot read RGBE image ") + path);
  const size_t w = img.GetWidth(), h = img.GetHeight();
  float *o = static_cast<float *>(std::malloc(w * h * 3 * sizeof(float)));
  if (!o) return io_fail(OMR_ERR_NOMEM, "out of host memory");
  for (size_t y = 0; y < h; y++)
    for (size_t x = 0; x < w; x++) {
      const OmochiRenderer::Color &c = img.GetPixel(x, y);
      float *p = o + (y * w + x) * 3;
      p[0] = (float)c.x;
      p[1] = (float)c.y;
      p[2] = (float)c.z;
    }
  *width = (uint32_t)w;
  *height = (uint32_t)h;
  *rgb = o;
  return OMR_OK;
}

extern "C" int omr_image_read(const char *path, uint32_t *width, uint32_t *height, uint32_t *channels,
                              uint8_t **px) {
  if (!path || !width || !height || !channels || !px) return io_fail(OMR_ERR_INVALID, "null argument");
  *px = nullptr;
  OmochiRenderer::codec::Pixels8 img;
  std::string err;
  if (!OmochiRenderer::codec::decode_file(path, img, err)) return io_fail(OMR_ERR_INVALID, std::string(path) + ": " + err);
  uint8_t *o = static_cast<uint8_t *>(std::malloc(img.px.size()));
  if (!o) return io_fail(OMR_ERR_NOMEM, "out of host memory");
  std::memcpy(o, img.px.data(), img.px.size());
  *width = img.width;
  *height = img.height;
  *channels = img.channels;
  *px = o;
  return OMR_OK;
}

extern "C" int omr_texture_read(const char *path, int inverse_gamma, uint32_t *width, uint32_t *height, float **rgb) {
  if (!path || !width || !height || !rgb) return io_fail(OMR_ERR_INVALID, "null argument");
  *rgb = nullptr;
  auto &ih = OmochiRenderer::ImageHandler::GetInstance();
  std::lock_guard<std::mutex> g(g_images_mu);
  const size_t before = ih.GetImageCount();
  OmochiRenderer::ImageHandler::IMAGE_ID id = ih.LoadFromFile(path, inverse_gamma != 0);
  if (id == OmochiRenderer::ImageHandler::INVALID_IMAGE_ID) return io_fail(OMR_ERR_INVALID, ih.LastError());
  const OmochiRenderer::Image *img = ih.GetImage(id);
  const size_t n = img->GetWidth() * img->GetHeight();
  float *o = static_cast<float *>(std::malloc(n * 3 * sizeof(float)));
  if (!o) return io_fail(OMR_ERR_NOMEM, "out of host memory");
  for (size_t i = 0; i < n; i++) {
    o[3 * i] = (float)img->m_image[i].x;
    o[3 * i + 1] = (float)img->m_image[i].y;
    o[3 * i + 2] = (float)img->m_image[i].z;
  }
  *width = (uint32_t)img->GetWidth();
  *height = (uint32_t)img->GetHeight();
  *rgb = o;
  if ((size_t)id >= before) ih.ReleaseImage(id);  // loaded for this call only
  return OMR_OK;
}

extern "C" int omr_hdr_write(const char *path, uint32_t width, uint32_t height, const float *rgb) {
  if (!path || !rgb || width == 0 || height == 0) return io_fail(OMR_ERR_INVALID, "bad argument");
  std::vector<OmochiRenderer::Color> px((size_t)width * height);
  for (size_t i = 0; i < px.size(); i++) px[i] = OmochiRenderer::Color(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
  OmochiRenderer::HDRImage img;
  img.SetImage(width, height, std::move(px));
  if (!img.WriteToRadianceFile(path)) return io_fail(OMR_ERR_INVALID, std::string("cannot write ") + path);
  return OMR_OK;
}

extern "C" int omr_scene_file_load(const char *path, omr_loaded_scene *out) {
  if (!path || !out) return io_fail(OMR_ERR_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  std::lock_guard<std::mutex> g(g_images_mu);
  std::unique_ptr<Loaded> L(new Loaded);
  auto *s = new OmochiRenderer::SceneFromExternalFile(path);
  L->scene.reset(s);
  if (!s->IsValid()) return io_fail(OMR_ERR_INVALID, std::string("cannot read scene file ") + path);
  return finish(L.release(), out);
}

extern "C" int omr_obj_load(const char *path, const double position[3], const double scale[3], const double rotation[3],
                            int flip_v, omr_loaded_scene *out) {
  if (!path || !out) return io_fail(OMR_ERR_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  using OmochiRenderer::Vector3;
  Vector3 pos = position ? Vector3(position[0], position[1], position[2]) : Vector3();
  Vector3 sc = scale ? Vector3(scale[0], scale[1], scale[2]) : Vector3(1, 1, 1);
  Vector3 rot = rotation ? Vector3(rotation[0], rotation[1], rotation[2]) : Vector3();
  std::lock_guard<std::mutex> g(g_images_mu);
  std::unique_ptr<Loaded> L(new Loaded);
  auto *s = new ObjScene;
  L->scene.reset(s);
  if (!s->Load(path, pos, sc, rot, flip_v != 0)) return io_fail(OMR_ERR_INVALID, std::string("cannot read OBJ ") + path);
  return finish(L.release(), out);
}

extern "C" void omr_loaded_scene_free(omr_loaded_scene *s) {
  if (!s) return;
  delete static_cast<Loaded *>(s->internal);
  std::memset(s, 0, sizeof(*s));
}
