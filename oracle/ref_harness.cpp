// ref_harness.cpp -- C entry points over the reference's own code, used only
// to write golden vectors in a container that has the reference checkout.
// It includes the reference headers from $(REF)/src and links the reference
// translation units listed in oracle/Makefile; nothing of the reference is
// copied here.  The output goes to oracle/_ref/ (never committed, never run
// on the GPU host).
#include "stdafx.h"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "renderer/IBL.h"
#include "renderer/InfiniteFloor.h"
#include "renderer/PathTracer.h"
#include "renderer/Polygon.h"
#include "renderer/Sphere.h"
#include "scenes/CornellBoxScene.h"
#include "scenes/Scene.h"
#include "scenes/SceneFromExternalFile.h"
#include "tools/HDRImage.h"
#include "tools/ImageHandler.h"
#include "stb/stb_image.h"
#include "tools/Random.h"

using namespace OmochiRenderer;

namespace {

struct Mat48 {
  uint32_t type;
  float color[3];
  float emission[3];
  float ior;
  float rate;
  uint32_t pad[3];
};

Material to_ref(const Mat48 &m) {
  return Material((Material::REFLECTION_TYPE)m.type, Vector3(m.emission[0], m.emission[1], m.emission[2]),
                  Vector3(m.color[0], m.color[1], m.color[2]), m.ior);
}

class HarnessScene : public Scene {
 public:
  std::vector<SceneObject *> prims;
  HarnessScene(const float *tri, const float *tri_n, size_t n_tri, const float *sph, size_t n_sph,
               const uint32_t *pm, const uint32_t *pc, const Mat48 *mats, int bvh) {
    auto add_mats = [&](SceneObject *o, size_t i) {
      uint32_t f = pm[i], c = pc ? pc[i] : 1u;
      for (uint32_t k = 1; k < c; k++)
        o->AddMaterial(to_ref(mats[f + k]), (double)mats[f + k].rate / (double)mats[f].rate, k + 1 == c);
    };
    for (size_t i = 0; i < n_tri; i++) {
      const float *t = tri + 9 * i, *n = tri_n + 9 * i;
      Vector3 p0(t[0], t[1], t[2]), e1(t[3], t[4], t[5]), e2(t[6], t[7], t[8]);
      Vector3 n0(n[0], n[1], n[2]), d1(n[3], n[4], n[5]), d2(n[6], n[7], n[8]);
      Polygon *p = new Polygon(p0, p0 + e1, p0 + e2, Vector3(), Vector3(), Vector3(), n0, n0 + d1, n0 + d2,
                               to_ref(mats[pm[i]]), Vector3(0, 0, 0));
      add_mats(p, i);
      AddObject(p);
      prims.push_back(p);
    }
    for (size_t i = 0; i < n_sph; i++) {
      const float *s = sph + 4 * i;
      Sphere *o = new Sphere(s[3], Vector3(s[0], s[1], s[2]), to_ref(mats[pm[n_tri + i]]));
      add_mats(o, n_tri + i);
      AddObject(o);
      prims.push_back(o);
    }
    if (bvh) ConstructBVH();
  }
  void SetEnv(const char *hdr) { m_ibl.reset(new IBL(hdr)); }
  // infinite floors outside the BVH, after every other object
  void AddPlanes(const float *y, size_t n, const uint32_t *pm, const uint32_t *pc, const Mat48 *mats, size_t first) {
    for (size_t i = 0; i < n; i++) {
      InfiniteFloor *o = new InfiniteFloor(y[i], to_ref(mats[pm[first + i]]));
      uint32_t f = pm[first + i], c = pc ? pc[first + i] : 1u;
      for (uint32_t k = 1; k < c; k++)
        o->AddMaterial(to_ref(mats[f + k]), (double)mats[f + k].rate / (double)mats[f].rate, k + 1 == c);
      AddObject(o, true, false);
      prims.push_back(o);
    }
  }
};

// A reference scene class with its object list readable: the .scene reader
// (SceneFromExternalFile) or the reference's own CornellBoxScene.
template <class Base>
class Listed : public Base {
 public:
  using Base::Base;
  size_t Count() const { return this->m_objects.size(); }
  size_t IndexOf(const SceneObject *o) const {
    for (size_t k = 0; k < this->m_objects.size(); k++)
      if (this->m_objects[k].object == o) return k;
    return (size_t)-1;
  }
};
using FileScene = Listed<SceneFromExternalFile>;
using CornellScene = Listed<CornellBoxScene>;

// Scene::CheckIntersection of a listed scene on n rays: FP64 distance,
// normal of the hit record, the object's material count, the object's index
// in the scene's object list (-1 on a miss), and its first four materials as
// floats (type, color, emission, ior, rate; 9 per material, 36 per ray).
template <class S>
void trace_listed(S *s, const float *rays, size_t n, double *t, double *normal, float *mats, int32_t *nmat,
                  int64_t *obj, double *uv) {
  for (size_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    Ray ray(Vector3(r[0], r[1], r[2]), Vector3(r[4], r[5], r[6]));
    Scene::IntersectionInformation info;
    bool hit = s->CheckIntersection(ray, info);
    obj[i] = -1;
    t[i] = info.hit.distance;
    if (!hit) continue;
    normal[3 * i] = info.hit.normal.x;
    normal[3 * i + 1] = info.hit.normal.y;
    normal[3 * i + 2] = info.hit.normal.z;
    uv[2 * i] = info.hit.uv.x;
    uv[2 * i + 1] = info.hit.uv.y;
    const size_t cnt = info.object->GetMaterialCount();
    for (size_t k = 0; k < cnt && k < 4; k++) {
      const SceneObject::MultiMaterialElem *e = info.object->GetMaterial(k);
      const Material &m = e->material_;
      float *o = mats + 36 * i + 9 * k;
      o[0] = (float)m.reflection_type;
      o[1] = (float)m.color.x; o[2] = (float)m.color.y; o[3] = (float)m.color.z;
      o[4] = (float)m.emission.x; o[5] = (float)m.emission.y; o[6] = (float)m.emission.z;
      o[7] = (float)m.refraction_rate;
      o[8] = (float)e->rate_;
    }
    nmat[i] = (int32_t)cnt;
    obj[i] = (int64_t)s->IndexOf(info.object);
  }
}

// Threads PathTracer::RenderScene runs with (SetMaxThreadCount; the
// reference's row bands write viewport threadIndex / 2, so 1 and 2 threads
// fill the one image), and the wall time of the last RenderScene.
size_t g_threads = 1;
double g_seconds = 0.0;

// PathTracer::RenderScene (src/renderer/PathTracer.cpp:55-85) on g_threads
// threads: `batches` sample batches of ss x ss samples each, pinhole camera.
// Writes the per-pixel mean and variance of the batch means (top-down rows,
// W x H x 3).
int render_scene(const Scene &scene, uint32_t W, uint32_t H, const double *pos, const double *dir, const double *up,
                 double screen_h, double dist, uint32_t ss, uint32_t batches, double *mean, double *var, int nee = 0) {
  Camera cam(W, H, Vector3(pos[0], pos[1], pos[2]), Vector3(dir[0], dir[1], dir[2]), Vector3(up[0], up[1], up[2]),
             screen_h, dist, 200.0);
  std::vector<double> prev((size_t)W * H * 3, 0.0), m1((size_t)W * H * 3, 0.0), m2((size_t)W * H * 3, 0.0);
  PathTracer::RenderingFinishCallbackFunction cb = [&](int samples, const Color *img, size_t w2, size_t h, double) {
    // left W x H half of the 2W x H buffer; stored row (H-1-y) is screen row y
    for (size_t r = 0; r < h; r++)
      for (size_t x = 0; x < W; x++) {
        const Color &c = img[x + r * w2];
        double cur[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
          size_t i = (r * W + x) * 3 + k;
          double batch = cur[k] * samples - prev[i] * (samples - 1);
          prev[i] = cur[k];
          m1[i] += batch;
          m2[i] += batch * batch;
        }
      }
  };
  // Heap-allocated and never destroyed: ~PathTracer applies delete[] to the
  // ScreenPixels it allocated with new (src/renderer/PathTracer.cpp:50-53,62),
  // which aborts under glibc.
  PathTracer *pt = new PathTracer(cam, 1, (int)batches, 1, (int)ss, cb);
  pt->SetMaxThreadCount(g_threads);
  pt->EnableNextEventEstimation(nee != 0);
  auto t0 = std::chrono::steady_clock::now();
  pt->RenderScene(scene);
  g_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (size_t i = 0; i < m1.size(); i++) {
    double mu = m1[i] / batches;
    mean[i] = mu;
    var[i] = batches > 1 ? (m2[i] - batches * mu * mu) / (batches - 1) : 0.0;
  }
  return 0;
}

}  // namespace

extern "C" {

// Random (src/tools/Random.h): n values of next() and, from a second
// generator with the same seed, n values of nextDouble().
// Thread count of the following renders (at most 2, see g_threads) and the
// wall time of the last one.
int ref_set_threads(size_t n) {
  if (n < 1 || n > 2) return -1;
  g_threads = n;
  return 0;
}
double ref_last_seconds() { return g_seconds; }

void ref_random(uint32_t seed, size_t n, uint32_t *out_next, double *out_double) {
  Random a(seed), b(seed);
  for (size_t i = 0; i < n; i++) {
    out_next[i] = a.next();
    out_double[i] = b.nextDouble();
  }
}

void *ref_scene_new(const float *tri, const float *tri_n, size_t n_tri, const float *sph, size_t n_sph,
                    const uint32_t *pm, const uint32_t *pc, const void *mats, int bvh, const char *hdr) {
  HarnessScene *s = new HarnessScene(tri, tri_n, n_tri, sph, n_sph, pm, pc, (const Mat48 *)mats, bvh);
  if (hdr && hdr[0]) s->SetEnv(hdr);
  return s;
}

void ref_scene_free(void *s) { delete (HarnessScene *)s; }

// Adds n infinite floors y[i] (materials pm/pc indexed from `first`).
void ref_scene_add_planes(void *s, const float *y, size_t n, const uint32_t *pm, const uint32_t *pc, const void *mats,
                          size_t first) {
  ((HarnessScene *)s)->AddPlanes(y, n, pm, pc, (const Mat48 *)mats, first);
}

// SceneFromExternalFile on `path` (relative names resolve against the
// current directory, as for the reference's executable).  NULL if invalid.
void *ref_file_scene_new(const char *path) {
  FileScene *s = new FileScene(path);
  if (!s->IsValid()) {
    delete s;
    return nullptr;
  }
  return s;
}
void ref_file_scene_free(void *s) { delete (FileScene *)s; }
// PathTracer::RenderScene over the loaded scene (see render_scene).
int ref_file_render(void *sp, uint32_t W, uint32_t H, const double *pos, const double *dir, const double *up,
                    double screen_h, double dist, uint32_t ss, uint32_t batches, double *mean, double *var) {
  return render_scene(*(FileScene *)sp, W, H, pos, dir, up, screen_h, dist, ss, batches, mean, var);
}
size_t ref_file_scene_objects(void *s) { return ((FileScene *)s)->Count(); }
// Scene::CheckIntersection of the loaded scene on n rays: FP64 distance,
// normal of the hit record, the object's material count, the object's index
// in the scene's object list (-1 on a miss), and its first four materials as
// floats (type, color, emission, ior, rate; 9 per material, 36 per ray).
void ref_file_trace(void *sp, const float *rays, size_t n, double *t, double *normal, float *mats, int32_t *nmat,
                    int64_t *obj, double *uv) {
  trace_listed((FileScene *)sp, rays, n, t, normal, mats, nmat, obj, uv);
}

// The reference's own CornellBoxScene (src/scenes/CornellBoxScene.cpp:10-36):
// floors, a two-material Lambert/mirror sphere, mirror and glass spheres and
// a SphereLight, no BVH (brute-force closest hit over its object list).
void *ref_cornell_new() { return new CornellScene(); }
void ref_cornell_free(void *s) { delete (CornellScene *)s; }
size_t ref_cornell_objects(void *s) { return ((CornellScene *)s)->Count(); }
void ref_cornell_trace(void *sp, const float *rays, size_t n, double *t, double *normal, float *mats, int32_t *nmat,
                       int64_t *obj, double *uv) {
  trace_listed((CornellScene *)sp, rays, n, t, normal, mats, nmat, obj, uv);
}
// nee: PathTracer::EnableNextEventEstimation (PathTracer.h:30-33)
int ref_cornell_render(void *sp, uint32_t W, uint32_t H, const double *pos, const double *dir, const double *up,
                       double screen_h, double dist, uint32_t ss, uint32_t batches, double *mean, double *var,
                       int nee) {
  return render_scene(*(CornellScene *)sp, W, H, pos, dir, up, screen_h, dist, ss, batches, mean, var, nee);
}

// HDRImage::ReadFromRadianceFile: dims into wh (w, h); pixels (h x w x 3,
// first scanline first) into out when out != NULL.
int ref_hdr_read(const char *path, int32_t *wh, float *out) {
  HDRImage img;
  if (!img.ReadFromRadianceFile(path)) return -1;
  wh[0] = (int32_t)img.GetWidth();
  wh[1] = (int32_t)img.GetHeight();
  if (out)
    for (size_t y = 0; y < img.GetHeight(); y++)
      for (size_t x = 0; x < img.GetWidth(); x++) {
        const Color &c = img.GetPixel(x, y);
        float *o = out + 3 * (y * img.GetWidth() + x);
        o[0] = (float)c.x; o[1] = (float)c.y; o[2] = (float)c.z;
      }
  return 0;
}

// The reference's image library call (stbi_load with native channels, as
// ImageHandler.cpp:54): dims and channel count into whc, 8-bit samples into
// out when out != NULL.
int ref_image_load(const char *path, int32_t *whc, uint8_t *out) {
  int w, h, c;
  unsigned char *px = stbi_load(path, &w, &h, &c, 0);
  if (!px) return -1;
  whc[0] = w;
  whc[1] = h;
  whc[2] = c;
  if (out) std::memcpy(out, px, (size_t)w * h * c);
  stbi_image_free(px);
  return 0;
}

// ImageHandler::LoadFromFile + GetImage: the texture as the renderer sees it
// (floats, rows from the top).
int ref_texture_load(const char *path, int inverse_gamma, int32_t *wh, float *out) {
  ImageHandler &ih = ImageHandler::GetInstance();
  ImageHandler::IMAGE_ID id = ih.LoadFromFile(path, inverse_gamma != 0);
  if (id == ImageHandler::INVALID_IMAGE_ID) return -1;
  const Image *img = ih.GetImage(id);
  wh[0] = (int32_t)img->GetWidth();
  wh[1] = (int32_t)img->GetHeight();
  if (out)
    for (size_t i = 0; i < img->GetWidth() * img->GetHeight(); i++) {
      const Color &c = img->m_image[i];
      out[3 * i] = (float)c.x;
      out[3 * i + 1] = (float)c.y;
      out[3 * i + 2] = (float)c.z;
    }
  return 0;
}

// Sphere::CheckIntersection's texture uv (Sphere.h:52-75) for unit normals:
// a unit sphere at the origin with a textured material, hit from outside
// along -n at n.
void ref_sphere_uv(const double *nrm, size_t n, double *uv) {
  Material m(Material::REFLECTION_TYPE_LAMBERT, Vector3(0, 0, 0), Vector3(1, 1, 1));
  m.texture_id = 0;
  Sphere s(1.0, Vector3(0, 0, 0), m);
  for (size_t i = 0; i < n; i++) {
    Vector3 d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    Ray ray(d * 3.0, d * -1.0);
    HitInformation hit;
    uv[2 * i] = uv[2 * i + 1] = -1.0;
    if (s.CheckIntersection(ray, hit)) {
      uv[2 * i] = hit.uv.x;
      uv[2 * i + 1] = hit.uv.y;
    }
  }
}

// Image::GetPixelByUV (Image.h:21-30) on a w x h image: texel column and
// row of each (u, v), read off the address of the returned pixel.
void ref_texel_xy(int32_t w, int32_t h, const double *uv, size_t n, int32_t *xy) {
  Image img;
  img.m_width = (size_t)w;
  img.m_height = (size_t)h;
  img.m_image.resize((size_t)w * h);
  for (size_t i = 0; i < n; i++) {
    long k = (long)(&img.GetPixelByUV(uv[2 * i], uv[2 * i + 1]) - img.m_image.data());
    xy[2 * i] = (int32_t)(k % w);
    xy[2 * i + 1] = (int32_t)(k / w);
  }
}

// HDRImage::WriteToRadianceFile of a w x h float image.
int ref_hdr_write(const char *path, int32_t w, int32_t h, const float *rgb) {
  HDRImage img;
  img.m_width = (size_t)w;
  img.m_height = (size_t)h;
  img.m_image.resize((size_t)w * h);
  for (size_t i = 0; i < img.m_image.size(); i++) img.m_image[i] = Color(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
  return img.WriteToRadianceFile(path) ? 0 : -1;
}

// Scene::CheckIntersection on n rays (float o.xyz, tmax, d.xyz, pad): the
// primitive index (-1 on a miss) and the FP64 distance.
void ref_trace(void *sp, const float *rays, size_t n, int32_t *prim, double *t) {
  HarnessScene *s = (HarnessScene *)sp;
  for (size_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    Ray ray(Vector3(r[0], r[1], r[2]), Vector3(r[4], r[5], r[6]));
    Scene::IntersectionInformation info;
    bool hit = s->CheckIntersection(ray, info);
    prim[i] = -1;
    t[i] = info.hit.distance;
    if (hit) {
      for (size_t k = 0; k < s->prims.size(); k++)
        if (s->prims[k] == info.object) { prim[i] = (int32_t)k; break; }
    }
  }
}

// IBL::Sample (src/renderer/IBL.h:30-45): texel of each direction, read off
// the address of the returned pixel.
int ref_env_texel(const char *hdr, size_t n, const double *dirs, int32_t *ix, int32_t *iy, int32_t *wh) {
  HDRImage img;
  if (!img.ReadFromRadianceFile(hdr)) return -1;
  IBL ibl(hdr);
  const double e = 1e-9;
  const Color *base = &ibl.Sample(Vector3(std::sin(e), std::cos(e), 0.0));
  int W = (int)img.GetWidth();
  wh[0] = W;
  wh[1] = (int)img.GetHeight();
  for (size_t i = 0; i < n; i++) {
    const Color *p = &ibl.Sample(Vector3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    long idx = (long)(p - base);
    ix[i] = (int32_t)(idx % W);
    iy[i] = (int32_t)(idx / W);
  }
  return 0;
}

// PathTracer::RenderScene over a harness scene (see render_scene).
int ref_render(void *sp, uint32_t W, uint32_t H, const double *pos, const double *dir, const double *up,
               double screen_h, double dist, uint32_t ss, uint32_t batches, double *mean, double *var) {
  return render_scene(*(HarnessScene *)sp, W, H, pos, dir, up, screen_h, dist, ss, batches, mean, var);
}

}  // extern "C"
