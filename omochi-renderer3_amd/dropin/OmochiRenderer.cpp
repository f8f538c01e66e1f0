// OmochiRenderer.cpp -- drop-in classes over the C ABI.
#include "OmochiRenderer.h"

#include <chrono>
#include <cstring>
#include <iostream>
#include <mutex>
#include <sstream>

namespace OmochiRenderer {

omr_ctx *DefaultContext() {
  static std::once_flag once;
  static omr_ctx *ctx = nullptr;
  std::call_once(once, [] {
    if (omr_create(&ctx, 0) != OMR_OK) ctx = nullptr;
  });
  return ctx;
}

void SceneObject::AddMaterial(const Material &m, double rate, bool doNormalizeRate) {
  materials_.emplace_back(m, rate);
  if (doNormalizeRate) {
    double total = 0;
    for (auto &e : materials_) total += e.rate_;
    for (auto &e : materials_) e.rate_ /= total;
  }
}

Polygon::Polygon(const Vector3 &pos1, const Vector3 &pos2, const Vector3 &pos3, const Vector3 &uv1, const Vector3 &uv2,
                 const Vector3 &uv3, const Vector3 &normal1, const Vector3 &normal2, const Vector3 &normal3,
                 const Material &mat, const Vector3 &pos)
    : SceneObject(mat) {
  m_posAndEdges[0] = pos1;
  m_posAndEdges[1] = pos2 - pos1;
  m_posAndEdges[2] = pos3 - pos1;
  m_uvOrigAndEdges[0] = uv1;
  m_uvOrigAndEdges[1] = uv2 - uv1;
  m_uvOrigAndEdges[2] = uv3 - uv1;
  m_normalAndDiffs[0] = normal1;
  m_normalAndDiffs[1] = normal2 - normal1;
  m_normalAndDiffs[2] = normal3 - normal1;
  position_ = pos;
}

void Polygon::Transform(const Vector3 &pos, const Vector3 &scale, const Matrix &rot) {
  position_ = pos;
  for (int i = 0; i < 3; i++) {
    m_posAndEdges[i] = rot.Apply(m_posAndEdges[i]);
    m_posAndEdges[i].x *= scale.x;
    m_posAndEdges[i].y *= scale.y;
    m_posAndEdges[i].z *= scale.z;
    m_normalAndDiffs[i] = rot.Apply(m_normalAndDiffs[i]);
  }
}

Vector3 Polygon::CalculateNormal(const Vector3 &v0, const Vector3 &v1, const Vector3 &v2) {
  Vector3 n((v1 - v0).cross(v2 - v0));
  n.normalize();
  return n;
}

bool Polygon::CheckIntersection(const Ray &ray, HitInformation &hit) const {
  const Vector3 &e1 = m_posAndEdges[1], &e2 = m_posAndEdges[2];
  const Vector3 P(ray.dir.cross(e2));
  const double det = P.dot(e1);
  if (!(det > EPS)) return false;  // single-sided
  const Vector3 T(ray.orig - (m_posAndEdges[0] + position_));
  const double u = P.dot(T);
  if (!(u >= 0 && u <= det)) return false;
  const Vector3 Q(T.cross(e1));
  const double v = Q.dot(ray.dir);
  if (!(v >= 0 && u + v <= det)) return false;
  const double t = Q.dot(e2) / det;
  if (!(t >= EPS)) return false;
  const double ur = u / det, vr = v / det;
  hit.distance = t;
  hit.position = ray.orig + ray.dir * t;
  hit.normal = m_normalAndDiffs[1] * ur + m_normalAndDiffs[2] * vr + m_normalAndDiffs[0];
  hit.normal.normalize();
  hit.uv = m_uvOrigAndEdges[1] * ur + m_uvOrigAndEdges[2] * vr + m_uvOrigAndEdges[0];
  return true;
}

bool Sphere::CheckIntersection(const Ray &ray, HitInformation &hit) const {
  const Vector3 &x = ray.orig, &v = ray.dir, &c = position_;
  const Vector3 cx = c - x;
  const double b = v.dot(cx);
  const double D = b * b - v.lengthSq() * (cx.lengthSq() - m_radius * m_radius);
  if (D < 0) return false;
  const double sD = std::sqrt(D), t1 = sD + b, t2 = -sD + b;
  if (t1 <= EPS && t2 <= EPS) return false;
  hit.distance = t2 > EPS ? t2 : t1;
  hit.position = x + v * hit.distance;
  hit.normal = hit.position - position_;
  hit.normal.normalize();
  bool textured = false;
  for (const auto &m : materials_) textured = textured || m.material_.texture_id != ImageHandler::INVALID_IMAGE_ID;
  if (textured) {  // latitude-longitude uv, origin at the top left of the image
    double tv = (std::asin(hit.normal.y) + PI * 0.5) / PI;
    while (tv > 1) tv -= 1;
    while (tv < 0) tv += 1;
    double tu = (std::atan2(hit.normal.z, hit.normal.x) + PI) / (2 * PI);
    while (tu > 1) tu -= 1;
    while (tu < 0) tu += 1;
    hit.uv.x = tu;
    hit.uv.y = 1 - tv;
  }
  return true;
}

void SphereLight::SampleOnePoint(Vector3 &point, Vector3 &normal, double &pdf, const Random &rnd) const {
  pdf = 1.0 / (4 * PI * m_radius * m_radius);
  const double phi = 2 * PI * rnd.nextDouble();
  const double ct = 1 - 2 * rnd.nextDouble(), st = std::sqrt(1 - ct * ct);
  const Vector3 dir(st * std::cos(phi), ct, st * std::sin(phi));
  point = position_ + dir * (m_radius + EPS);
  normal = dir;
}

bool SphereLight::SampleOnePointWithTargetPoint(Vector3 &sampledPoint, Vector3 &sampledPointNormal, double &pdf,
                                                const Vector3 &targetPoint, const Random &rnd) const {
  Vector3 axis = targetPoint - position_;
  const double dist = axis.length();
  const double max_ct = m_radius / dist;
  if (max_ct > 1) return false;  // the target is inside the sphere
  axis /= dist;
  Vector3 a1, a2;
  Utils::GetCrossAxes(axis, a1, a2);
  pdf = 1.0 / (2 * PI * m_radius * m_radius * (1 - max_ct));
  const double phi = 2 * PI * rnd.nextDouble();
  const double ct = 1 - (1 - max_ct) * rnd.nextDouble(), st = std::sqrt(1 - ct * ct);
  const Vector3 dir = axis * ct + a1 * st * std::sin(phi) + a2 * st * std::cos(phi);
  sampledPoint = position_ + dir * (m_radius + EPS);
  sampledPointNormal = dir;
  return true;
}

double SphereLight::TotalPower() const {
  double e = 0;
  for (const auto &m : materials_) e += m.material_.emission.length() * m.rate_;
  return 4 * PI * m_radius * m_radius * e;
}

bool InfiniteFloor::CheckIntersection(const Ray &ray, HitInformation &hit) const {
  if (std::fabs(ray.dir.y) < EPS) return false;
  const double t = (m_y - ray.orig.y) / ray.dir.y;
  if (t < EPS) return false;
  hit.distance = t;
  hit.position = ray.orig + ray.dir * t;
  hit.normal = Vector3(0, ray.dir.y < 0 ? 1 : -1, 0);
  return true;
}

Camera::Camera(size_t w, size_t h, const Vector3 &pos, const Vector3 &dir, const Vector3 &up, double screenH,
               double dist, double focal)
    : m_width(w), m_height(h), m_camPos(pos), m_camDir(dir), m_camUp(up), m_nearScreenHeight(screenH),
      m_distToScreen(dist), m_focalDistance(focal) {
  // Camera::Camera (src/renderer/Camera.h:12-43)
  m_camDir.normalize();
  m_camUp.normalize();
  double nearW = m_nearScreenHeight * (double)m_width / (double)m_height;
  m_screenCenter = m_camPos + m_camDir * m_distToScreen;
  m_screenXaxis = m_camDir.cross(m_camUp);
  m_screenXaxis.normalize();
  m_screenXaxis *= nearW;
  m_screenYaxis = m_screenXaxis.cross(m_camDir);
  m_screenYaxis.normalize();
  m_screenYaxis *= m_nearScreenHeight;
}

Ray Camera::SampleRayForPixel(double x, double y, const Random &rnd) const {
  Vector3 target = m_screenCenter + m_screenXaxis * (x / m_width - 0.5) + m_screenYaxis * (y / m_height - 0.5);
  Vector3 dir = target - m_camPos;
  dir.normalize();
  Ray ray(m_camPos, dir);
  if (m_aperture) {
    double dx, dy;
    m_aperture->SampleOnePoint(dx, dy, rnd);
    const Vector3 focus = ray(m_focalDistance / std::fabs(dir.z));  // focal plane of the centre ray
    ray.orig = m_camPos + Vector3(dx, dy, 0);
    ray.dir = (focus - ray.orig).normalize();
  }
  return ray;
}

omr_camera Camera::ToDevice() const {
  omr_camera c;
  std::memset(&c, 0, sizeof(c));
  c.width = (uint32_t)m_width;
  c.height = (uint32_t)m_height;
  const Vector3 *v[4] = {&m_camPos, &m_screenCenter, &m_screenXaxis, &m_screenYaxis};
  float *dst[4] = {c.position, c.screen_center, c.screen_x, c.screen_y};
  for (int i = 0; i < 4; i++) {
    dst[i][0] = (float)v[i]->x;
    dst[i][1] = (float)v[i]->y;
    dst[i][2] = (float)v[i]->z;
  }
  c.aperture_radius = m_aperture ? (m_aperture->Radius() < 0 ? -1.0f : (float)m_aperture->Radius()) : 0.0f;
  c.focal_distance = (float)m_focalDistance;
  return c;
}

Scene::~Scene() {
  if (m_ctx && m_id >= 0) omr_scene_release(m_ctx, m_id);
  for (auto &o : m_objects)
    if (o.doDelete) delete o.obj;
}

void Scene::AddObject(SceneObject *obj, bool doDelete, bool) {
  m_objects.push_back({obj, doDelete});
  if (LightBase *l = dynamic_cast<LightBase *>(obj)) m_lights.push_back(l);
  if (m_ctx && m_id >= 0) {  // contents changed: re-upload on next use
    omr_scene_release(m_ctx, m_id);
    m_id = -1;
  }
}

void Scene::ConstructQBVH() {
  omr_ctx *c = DefaultContext();
  if (c) Upload(c);
}

void Scene::AddModel(Model *model, bool doDelete, bool containedInBVH) {
  // every polygon, material by material (Scene.h:78-88); a material of the
  // same value as an earlier one shares its list and is added once
  std::vector<const Model::PolygonList *> seen;
  for (size_t i = 0; i < model->GetMaterialCount(); i++) {
    const Model::PolygonList &pl = model->GetPolygonList(model->GetMaterial(i));
    bool dup = false;
    for (auto *q : seen) dup = dup || q == &pl;
    if (dup) continue;
    seen.push_back(&pl);
    for (Polygon *p : pl) AddObject(p, false, containedInBVH);
  }
  if (doDelete) m_models.emplace_back(model);
}

SceneObject *Scene::AddInfiniteFLoor(double y, const Material &material) {
  SceneObject *o = new InfiniteFloor(y, material);
  AddObject(o, true, false);
  return o;
}

void Scene::Describe(SceneArrays &a) const {
  a = SceneArrays();
  std::vector<uint32_t> pm_sph, pc_sph, pm_pln, pc_pln;
  std::vector<SceneObject *> sphs, plns;
  // images used as textures, numbered in order of first use (texture = index + 1)
  std::map<ImageHandler::IMAGE_ID, uint32_t> tex_index;
  auto texture_of = [&](ImageHandler::IMAGE_ID id) -> uint32_t {
    if (id == ImageHandler::INVALID_IMAGE_ID) return 0u;
    auto it = tex_index.find(id);
    if (it != tex_index.end()) return it->second;
    const Image *img = ImageHandler::GetInstance().GetImage(id);
    if (!img || img->GetWidth() == 0 || img->GetHeight() == 0) return 0u;
    std::vector<float> rgb(img->m_image.size() * 3);
    for (size_t i = 0; i < img->m_image.size(); i++) {
      rgb[3 * i] = (float)img->m_image[i].x;
      rgb[3 * i + 1] = (float)img->m_image[i].y;
      rgb[3 * i + 2] = (float)img->m_image[i].z;
    }
    a.tex_rgb.push_back(std::move(rgb));
    omr_texture t;
    t.width = (uint32_t)img->GetWidth();
    t.height = (uint32_t)img->GetHeight();
    t.rgb = nullptr;  // set once tex_rgb stops growing
    a.textures.push_back(t);
    return tex_index[id] = (uint32_t)a.textures.size();
  };
  auto add_mats = [&](const SceneObject *o, std::vector<uint32_t> &pm, std::vector<uint32_t> &pc) {
    pm.push_back((uint32_t)a.mats.size());
    pc.push_back((uint32_t)o->GetMaterialCount());
    for (size_t i = 0; i < o->GetMaterialCount(); i++) {
      const SceneObject::MultiMaterialElem *e = o->GetMaterial(i);
      omr_material m;
      std::memset(&m, 0, sizeof(m));
      m.type = (uint32_t)e->material_.reflection_type;
      const Color &col = e->material_.color, &em = e->material_.emission;
      m.color[0] = (float)col.x; m.color[1] = (float)col.y; m.color[2] = (float)col.z;
      m.emission[0] = (float)em.x; m.emission[1] = (float)em.y; m.emission[2] = (float)em.z;
      m.ior = (float)e->material_.refraction_rate;
      m.rate = (float)e->rate_;
      m.texture = texture_of(e->material_.texture_id);
      a.mats.push_back(m);
    }
  };
  for (auto &ow : m_objects) {
    SceneObject *o = ow.obj;
    if (o->GetMaterialCount() == 0) continue;
    if (auto *p = dynamic_cast<Polygon *>(o)) {
      const Vector3 *pe = p->PosAndEdges(), *nd = p->NormalAndDiffs();
      Vector3 p0 = pe[0] + p->GetPosition();  // Polygon.h:158: T = o - (p0 + position_)
      float t9[9] = {(float)p0.x, (float)p0.y, (float)p0.z, (float)pe[1].x, (float)pe[1].y, (float)pe[1].z,
                     (float)pe[2].x, (float)pe[2].y, (float)pe[2].z};
      float n9[9] = {(float)nd[0].x, (float)nd[0].y, (float)nd[0].z, (float)nd[1].x, (float)nd[1].y, (float)nd[1].z,
                     (float)nd[2].x, (float)nd[2].y, (float)nd[2].z};
      const Vector3 *uv = p->UVOrigAndEdges();
      float u6[6] = {(float)uv[0].x, (float)uv[0].y, (float)uv[1].x, (float)uv[1].y, (float)uv[2].x, (float)uv[2].y};
      a.tri.insert(a.tri.end(), t9, t9 + 9);
      a.nrm.insert(a.nrm.end(), n9, n9 + 9);
      a.uv.insert(a.uv.end(), u6, u6 + 6);
      add_mats(o, a.prim_mat, a.prim_mat_count);
      a.prims.push_back(o);
    } else if (auto *s = dynamic_cast<Sphere *>(o)) {
      const Vector3 &c = s->GetPosition();
      float s4[4] = {(float)c.x, (float)c.y, (float)c.z, (float)s->GetRadius()};
      a.sph.insert(a.sph.end(), s4, s4 + 4);
      add_mats(o, pm_sph, pc_sph);
      sphs.push_back(o);
    } else if (auto *f = dynamic_cast<InfiniteFloor *>(o)) {
      a.planes.push_back((float)f->GetY());
      add_mats(o, pm_pln, pc_pln);
      plns.push_back(o);
    }
  }
  a.prim_mat.insert(a.prim_mat.end(), pm_sph.begin(), pm_sph.end());
  a.prim_mat_count.insert(a.prim_mat_count.end(), pc_sph.begin(), pc_sph.end());
  a.prim_mat.insert(a.prim_mat.end(), pm_pln.begin(), pm_pln.end());
  a.prim_mat_count.insert(a.prim_mat_count.end(), pc_pln.begin(), pc_pln.end());
  a.prims.insert(a.prims.end(), sphs.begin(), sphs.end());
  a.prims.insert(a.prims.end(), plns.begin(), plns.end());
  omr_scene_desc &d = a.desc;
  std::memset(&d, 0, sizeof(d));
  d.n_triangles = a.tri.size() / 9;
  d.triangles = a.tri.data();
  d.tri_normals = a.nrm.data();
  d.n_spheres = a.sph.size() / 4;
  d.spheres = a.sph.data();
  d.n_planes = a.planes.size();
  d.planes = a.planes.data();
  d.prim_material = a.prim_mat.data();
  d.prim_material_count = a.prim_mat_count.data();
  d.n_materials = a.mats.size();
  d.materials = a.mats.data();
  for (size_t k = 0; k < a.textures.size(); k++) a.textures[k].rgb = a.tex_rgb[k].data();
  d.n_textures = a.textures.size();
  d.textures = a.textures.empty() ? nullptr : a.textures.data();
  d.tri_uv = a.textures.empty() ? nullptr : a.uv.data();  // uv only matters with a texture
  if (m_ibl && m_ibl->IsValid()) {
    d.env_width = m_ibl->Width();
    d.env_height = m_ibl->Height();
    d.env_rgb = m_ibl->Pixels().data();
  }
}

omr_scene_id Scene::Upload(omr_ctx *ctx) const {
  if (m_ctx == ctx && m_id >= 0) return m_id;
  SceneArrays a;
  Describe(a);
  omr_scene_id id = -1;
  if (omr_scene_upload(ctx, &a.desc, &id) != OMR_OK) {
    std::cerr << "scene upload failed: " << omr_last_error(ctx) << std::endl;
    return -1;
  }
  m_prims = a.prims;
  m_ctx = ctx;
  m_id = id;
  return id;
}

bool Scene::CheckIntersection(const Ray &ray, IntersectionInformation &info) const {
  info.hit = HitInformation();
  info.hit.distance = INF;
  info.object = nullptr;
  omr_ctx *c = m_ctx ? m_ctx : DefaultContext();
  if (c && Upload(c) >= 0) {
    float r[8] = {(float)ray.orig.x, (float)ray.orig.y, (float)ray.orig.z, 0x1.fffffep+126f,
                  (float)ray.dir.x, (float)ray.dir.y, (float)ray.dir.z, 0.0f};
    int32_t prim = -1;
    float t = 0;
    if (omr_trace_closest(c, m_id, r, 1, &prim, &t) == OMR_OK) {
      if (prim < 0) return false;
      // the object the GPU found computes its own FP64 hit record
      SceneObject *o = m_prims[(size_t)prim];
      if (o->CheckIntersection(ray, info.hit)) {
        info.object = o;
        return true;
      }
    }
  }
  // no device, or a hit the FP64 test rejects (a grazing binary32 hit): test
  // every object in FP64, the first of equal distances winning as in
  // Scene.cpp:51-85
  for (const auto &ow : m_objects) {
    HitInformation h;
    if (ow.obj->GetMaterialCount() && ow.obj->CheckIntersection(ray, h) && h.distance < info.hit.distance) {
      info.hit = h;
      info.object = ow.obj;
    }
  }
  return info.object != nullptr;
}

// Floor helpers: two polygons per quad, same vertices, uvs and normals as
// Scene::AddFloor* (src/scenes/Scene.cpp:88-193).
#define OMR_FLOOR(name, a0, a1, a2, ta, n0, b0, b1, b2, tb)                                                 \
  std::vector<SceneObject *> Scene::name(double s1, double s2, const Vector3 &position, const Material &m) { \
    const double h1 = s1 / 2, h2 = s2 / 2;                                                                   \
    (void)h1; (void)h2;                                                                                      \
    const double ua[6] = {ta}, ub[6] = {tb};                                                                 \
    std::vector<SceneObject *> objs;                                                                         \
    objs.push_back(new Polygon(a0, a1, a2, Vector3(ua[0], ua[1], 0), Vector3(ua[2], ua[3], 0),              \
                               Vector3(ua[4], ua[5], 0), n0, n0, n0, m, position));                          \
    objs.push_back(new Polygon(b0, b1, b2, Vector3(ub[0], ub[1], 0), Vector3(ub[2], ub[3], 0),              \
                               Vector3(ub[4], ub[5], 0), n0, n0, n0, m, position));                          \
    AddObject(objs[0]);                                                                                      \
    AddObject(objs[1]);                                                                                      \
    return objs;                                                                                             \
  }
#define OMR_UV(a, b, c, d, e, f) a, b, c, d, e, f
OMR_FLOOR(AddFloorXZ_yUp, Vector3(-h1, 0, -h2), Vector3(-h1, 0, h2), Vector3(h1, 0, -h2), OMR_UV(0, 0, 0, 1, 1, 0),
          Vector3(0, 1, 0), Vector3(h1, 0, -h2), Vector3(-h1, 0, h2), Vector3(h1, 0, h2), OMR_UV(1, 0, 0, 1, 1, 1))
OMR_FLOOR(AddFloorXY_zUp, Vector3(-h1, h2, 0), Vector3(-h1, -h2, 0), Vector3(h1, -h2, 0), OMR_UV(0, 0, 0, 1, 1, 1),
          Vector3(0, 0, 1), Vector3(h1, h2, 0), Vector3(-h1, h2, 0), Vector3(h1, -h2, 0), OMR_UV(1, 0, 0, 0, 1, 1))
OMR_FLOOR(AddFloorYZ_xUp, Vector3(0, h1, -h2), Vector3(0, h1, h2), Vector3(0, -h1, h2), OMR_UV(1, 0, 0, 0, 0, 1),
          Vector3(1, 0, 0), Vector3(0, -h1, -h2), Vector3(0, h1, -h2), Vector3(0, -h1, h2), OMR_UV(1, 1, 1, 0, 0, 1))
OMR_FLOOR(AddFloorXZ_yDown, Vector3(h1, 0, -h2), Vector3(-h1, 0, h2), Vector3(-h1, 0, -h2), OMR_UV(1, 1, 0, 0, 0, 1),
          Vector3(0, -1, 0), Vector3(h1, 0, h2), Vector3(-h1, 0, h2), Vector3(h1, 0, -h2), OMR_UV(1, 0, 0, 0, 1, 1))
OMR_FLOOR(AddFloorXY_zDown, Vector3(h1, -h2, 0), Vector3(-h1, -h2, 0), Vector3(-h1, h2, 0), OMR_UV(0, 1, 1, 1, 1, 0),
          Vector3(0, 0, -1), Vector3(h1, -h2, 0), Vector3(-h1, h2, 0), Vector3(h1, h2, 0), OMR_UV(0, 1, 1, 0, 0, 0))
OMR_FLOOR(AddFloorYZ_xDown, Vector3(0, -h1, h2), Vector3(0, h1, h2), Vector3(0, h1, -h2), OMR_UV(1, 1, 1, 0, 0, 0),
          Vector3(-1, 0, 0), Vector3(0, -h1, h2), Vector3(0, h1, -h2), Vector3(0, -h1, -h2), OMR_UV(1, 1, 0, 0, 0, 1))
#undef OMR_UV
#undef OMR_FLOOR

PathTracer::PathTracer(const Camera &camera, int samples, int supersamples)
    : m_camera(camera), m_min(samples), m_max(samples), m_step(1), m_ss(supersamples) {}

PathTracer::PathTracer(const Camera &camera, int min_samples, int max_samples, int step, int supersamples,
                       RenderingFinishCallbackFunction callback)
    : m_camera(camera), m_min(min_samples), m_max(max_samples), m_step(step), m_ss(supersamples),
      m_callback(callback) {}

PathTracer::~PathTracer() {}

// PathTracer::RenderScene (src/renderer/PathTracer.cpp:55-85): progressive
// sample batches, running mean img*prev/next + batch_mean*(next-prev)/next.
void PathTracer::RenderScene(const Scene &scene) {
  m_error.clear();
  const size_t W = m_camera.GetScreenWidth(), H = m_camera.GetScreenHeight();
  m_result.reset(new ScreenPixels());
  m_result->InitializePixels(W, H, 2, 1);  // PathTracer.cpp:62-63
  m_ctx = DefaultContext();
  if (!m_ctx) { m_error = "no GPU context"; return; }
  omr_scene_id id = scene.Upload(m_ctx);
  if (id < 0) { m_error = omr_last_error(m_ctx); return; }
  omr_camera cam = m_camera.ToDevice();
  if (cam.aperture_radius < 0) { m_error = "only circular apertures (CircleAperture) render on the GPU"; return; }
  std::vector<float> batch(W * H * 3);
  m_previousSamples = 0;
  double minutes = 0;
  for (m_currentSamples = m_min; m_currentSamples <= m_max && m_enableRendering; m_currentSamples += m_step) {
    auto t0 = std::chrono::steady_clock::now();
    omr_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.supersamples = (uint32_t)m_ss;
    p.spp_begin = (uint32_t)m_previousSamples;
    p.spp_end = (uint32_t)m_currentSamples;
    p.max_depth = m_maxDepth;
    p.rr_depth = m_rrDepth;
    p.env_is = m_envIS ? 1u : 0u;
    p.nee = m_nee ? 1u : 0u;
    p.seed = m_seed;
    if (omr_render(m_ctx, id, &cam, &p, batch.data()) != OMR_OK) {
      m_error = omr_last_error(m_ctx);
      return;
    }
    omr_render_stats st;
    omr_last_render_stats(m_ctx, &st);
    m_segments += st.segments;
    double prev = m_previousSamples, next = m_currentSamples;
    Color *px = m_result->GetScreen();
    for (size_t y = 0; y < H; y++) {
      // output row r = H-1-y holds screen row y; ScreenPixels stores y at H-y-1
      const float *row = &batch[(H - 1 - y) * W * 3];
      for (size_t x = 0; x < W; x++) {
        Color &c = px[x + (H - y - 1) * 2 * W];
        Color b(row[3 * x], row[3 * x + 1], row[3 * x + 2]);
        c = c * (prev / next) + b * ((next - prev) / next);
      }
    }
    m_previousSamples = m_currentSamples;
    minutes += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
    if (m_callback) m_callback(m_currentSamples, m_result->GetScreen(), 2 * W, H, minutes);
  }
}

std::string PathTracer::GetCurrentRenderingInfo() const {
  std::stringstream ss;
  ss << "(width, height) = (" << m_camera.GetScreenWidth() << ", " << m_camera.GetScreenHeight() << ")" << std::endl;
  ss << "previous samples / pixel = " << m_previousSamples << "x(" << m_ss << "x" << m_ss << ")" << std::endl;
  ss << "segments traced = " << m_segments << std::endl;
  return ss.str();
}

}  // namespace OmochiRenderer
