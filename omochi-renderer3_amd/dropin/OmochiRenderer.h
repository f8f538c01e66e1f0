// OmochiRenderer.h -- C++ drop-in for the reference's renderer interface,
// backed by the MI355X library through include/omr_capi.h.
//
// Same names, argument meaning and error behaviour as the reference types
// (paths relative to the reference checkout):
//   Vector3 / Color        src/tools/Vector.h:9-103, src/renderer/Color.h
//   Ray                    src/renderer/Ray.h:7-20
//   Material               src/renderer/Material.h:11-36
//   SceneObject            src/renderer/SceneObject.h:17-95 (multi-material rates)
//   Sphere, Polygon        src/renderer/Sphere.h:11-82, src/renderer/Polygon.h:11-125
//   SphereLight, LightBase src/renderer/SphereLight.h:8-79, src/renderer/LightBase.h (light sampling)
//   HitInformation         src/renderer/HitInformation.h:7-20
//   Random                 src/tools/Random.h:9-29
//   Utils::GetCrossAxes    src/tools/Utils.h:123-131
//   INF, PI, EPS, ...      src/tools/Constant.h:4-8
//   Scene                  src/scenes/Scene.h:18-124 (AddObject, AddFloor*, ConstructQBVH)
//   Camera, CircleAperture src/renderer/Camera.h:10-117, src/renderer/Aperture.h:9-39
//   ScreenPixels           src/renderer/ScreenPixels.hpp / .cpp (2W x H, row-flipped)
//   Renderer, PathTracer   src/renderer/Renderer.h:10-23, src/renderer/PathTracer.h:18-88
//   SceneFactory(Manager)  src/scenes/SceneFactory.h:7-48
//   Matrix                 src/tools/Matrix.h:8-65 (rotations for mesh placement)
//   HDRImage               src/tools/HDRImage.h:8-81, HDRImage.cpp:12-314 (Radiance RGBE)
//   ImageHandler           src/tools/ImageHandler.h (texture registry; ids only)
//   InfiniteFloor          src/renderer/InfiniteFloor.h:10-45
//   Model                  src/renderer/Model.h:13-63, Model.cpp:14-367 (Wavefront OBJ/MTL)
//   SceneFromExternalFile  src/scenes/SceneFromExternalFile.h:14-52, .cpp:13-531 (.scene)
// The path tracing itself runs on the GPU; Scene::CheckIntersection runs the
// device closest-hit kernel for single rays (slow; for compatibility).
#pragma once
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/omr_capi.h"

// Constants of the reference (src/tools/Constant.h:4-8), as macros as there.
#ifndef INF
#define INF ((FLT_MAX - 1) / 2)
#endif
#ifndef PI
#define PI 3.14159265358979323846264338327950288419716939
#endif
#ifndef EPS
#define EPS 0.0005
#endif
#ifndef REFRACTIVE_INDEX_VACUUM
#define REFRACTIVE_INDEX_VACUUM 1.0
#endif
#ifndef REFRACTIVE_INDEX_OBJECT
#define REFRACTIVE_INDEX_OBJECT 1.5
#endif

namespace OmochiRenderer {

class Vector3 {
 public:
  double x, y, z;
  Vector3() : x(0), y(0), z(0) {}
  Vector3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
  Vector3 &operator+=(const Vector3 &r) { x += r.x; y += r.y; z += r.z; return *this; }
  Vector3 &operator-=(const Vector3 &r) { x -= r.x; y -= r.y; z -= r.z; return *this; }
  Vector3 &operator*=(double v) { x *= v; y *= v; z *= v; return *this; }
  Vector3 &operator/=(double v) { x /= v; y /= v; z /= v; return *this; }
  bool operator==(const Vector3 &v) const { return x == v.x && y == v.y && z == v.z; }
  Vector3 operator*(double v) const { return Vector3(x * v, y * v, z * v); }
  Vector3 operator/(double v) const { return Vector3(x / v, y / v, z / v); }
  Vector3 operator+(const Vector3 &v) const { return Vector3(x + v.x, y + v.y, z + v.z); }
  Vector3 operator-(const Vector3 &v) const { return Vector3(x - v.x, y - v.y, z - v.z); }
  double dot(const Vector3 &r) const { return x * r.x + y * r.y + z * r.z; }
  Vector3 cross(const Vector3 &r) const { return Vector3(y * r.z - z * r.y, z * r.x - x * r.z, x * r.y - y * r.x); }
  double lengthSq() const { return x * x + y * y + z * z; }
  double length() const { return std::sqrt(lengthSq()); }
  Vector3 &normalize() { return (*this) /= length(); }
  static const Vector3 &Zero() { static Vector3 z; return z; }
  static const Vector3 &One() { static Vector3 o(1, 1, 1); return o; }
};
typedef Vector3 Color;

class Ray {
 public:
  Ray(const Vector3 &o, const Vector3 &d) : orig(o), dir(d) {}
  Vector3 operator()(double t) const { return orig + dir * t; }
  Vector3 orig, dir;
};

// Random (src/tools/Random.h:9-29): xorshift128 with the reference's seeding
// recurrence; nextDouble() in [0, 1].
class Random {
 public:
  explicit Random(unsigned int seed) {
    unsigned int s = seed;
    for (unsigned int i = 1; i <= 4; i++) m_seed[i - 1] = s = 1812433253U * (s ^ (s >> 30)) + i;
  }
  unsigned int next() const {
    const unsigned int t = m_seed[0] ^ (m_seed[0] << 11);
    m_seed[0] = m_seed[1];
    m_seed[1] = m_seed[2];
    m_seed[2] = m_seed[3];
    return m_seed[3] = (m_seed[3] ^ (m_seed[3] >> 19)) ^ (t ^ (t >> 8));
  }
  double nextDouble() const { return static_cast<double>(next()) / UINT_MAX; }

 private:
  mutable unsigned int m_seed[4];
};

// Utils::GetCrossAxes (src/tools/Utils.h:123-131): the two axes the
// reference builds around a normal (unnormalized, as there).
class Utils {
 public:
  static void GetCrossAxes(const Vector3 &up, Vector3 &axis1, Vector3 &axis2) {
    axis1 = std::fabs(up.x) > EPS ? Vector3(0, 1, 0).cross(up) : Vector3(1, 0, 0).cross(up);
    axis2 = up.cross(axis1);
  }
};

// HitInformation (src/renderer/HitInformation.h:7-20).
class HitInformation {
 public:
  double distance = 0;
  Vector3 position;
  Vector3 normal;
  Vector3 uv;
};

// Texture registry (ImageHandler, src/tools/ImageHandler.h): file name -> id.
// Ids only: texture lookups are not part of the shading path.
// Image (src/tools/Image.h:4-35): pixels as Color, row 0 = top of the image.
class Image {
 public:
  virtual ~Image() {}
  size_t GetHeight() const { return m_height; }
  size_t GetWidth() const { return m_width; }
  const Color &GetPixel(size_t x, size_t y) const { return m_image[x + y * m_width]; }
  // u, v clamped to [0, 1] and scaled by 0.99999, nearest texel (Image.h:21-30)
  const Color &GetPixelByUV(double u, double v) const;

  std::vector<Color> m_image;
  size_t m_width = 0, m_height = 0;
};

// ImageHandler (src/tools/ImageHandler.h:8-39, ImageHandler.cpp:20-187): the
// process-wide image table that Material::texture_id indexes.  LoadFromFile
// reads .hdr files as Radiance RGBE (linear) and other files as 8-bit PNG or
// baseline JPEG, scaled to [0, 1] and, by default, raised to the power 2.2;
// the same LDR file name gives the same id.  INVALID_IMAGE_ID when the file
// cannot be read (the reason: LastError()).
class ImageHandler {
 public:
  typedef int IMAGE_ID;
  static const IMAGE_ID INVALID_IMAGE_ID = -1;
  static ImageHandler &GetInstance();
  IMAGE_ID LoadFromFile(const std::string &fname, bool doReverseGamma = true);
  IMAGE_ID CreateImage(size_t width, size_t height);
  Image *GetImage(IMAGE_ID id);
  const Image *GetImage(IMAGE_ID id) const;
  void ReleaseImage(IMAGE_ID id);
  const std::string &GetFileName(IMAGE_ID id) const;
  size_t GetImageCount() const { return m_images.size(); }
  const std::string &LastError() const { return m_error; }

 private:
  std::vector<std::unique_ptr<Image>> m_images;
  std::vector<std::string> m_files;  // parallel to m_images ("" for created images)
  std::string m_error;
};

class Material {
 public:
  enum REFLECTION_TYPE { REFLECTION_TYPE_LAMBERT, REFLECTION_TYPE_SPECULAR, REFLECTION_TYPE_REFRACTION };
  Material(REFLECTION_TYPE type = REFLECTION_TYPE_LAMBERT, const Vector3 &emission_ = Vector3(),
           const Vector3 &color_ = Vector3(), double refraction_rate_ = 0.0,
           ImageHandler::IMAGE_ID texture_id_ = ImageHandler::INVALID_IMAGE_ID)
      : reflection_type(type), emission(emission_), color(color_), refraction_rate(refraction_rate_),
        texture_id(texture_id_) {}
  bool operator==(const Material &m) const {
    return reflection_type == m.reflection_type && emission == m.emission && color == m.color &&
           refraction_rate == m.refraction_rate && texture_id == m.texture_id;
  }
  REFLECTION_TYPE reflection_type;
  Color emission;
  Color color;
  double refraction_rate;
  ImageHandler::IMAGE_ID texture_id;
};

// 4x4 matrix of which the upper 3x3 rotates (Matrix, src/tools/Matrix.h:8-65).
class Matrix {
 public:
  Matrix();
  explicit Matrix(const double m_[4][4]);
  static const Matrix &Identity();
  // rotation by `angle` radians about `axis` (normalized first)
  static Matrix RotateAroundVector(const Vector3 &axis, double angle);
  Vector3 Apply(const Vector3 &v) const;
  Matrix operator*(const Matrix &rhs) const;
  double m[4][4];
};

class SceneObject {
 public:
  struct MultiMaterialElem {
    Material material_;
    double rate_;
    MultiMaterialElem(const Material &m, double r = 1.0) : material_(m), rate_(r) {}
  };
  explicit SceneObject(const Material &m) { AddMaterial(m, 1.0, false); }
  SceneObject() {}
  virtual ~SceneObject() {}
  // SceneObject::CheckIntersection (src/renderer/SceneObject.h:46), in FP64
  // on the host as the reference computes it; rendering uses the GPU path.
  virtual bool CheckIntersection(const Ray &ray, HitInformation &hit) const = 0;
  // SceneObject::AddMaterial (src/renderer/SceneObject.h:48-67): rates renormalized
  void AddMaterial(const Material &m, double rate, bool doNormalizeRate = true);
  size_t GetMaterialCount() const { return materials_.size(); }
  const MultiMaterialElem *GetMaterial(size_t i) const { return i < materials_.size() ? &materials_[i] : nullptr; }
  MultiMaterialElem *GetMaterial(size_t i) { return i < materials_.size() ? &materials_[i] : nullptr; }
  const Vector3 &GetPosition() const { return position_; }

 protected:
  std::vector<MultiMaterialElem> materials_;
  Vector3 position_;
  friend class Scene;
};

class Sphere : public SceneObject {
 public:
  Sphere(double r, const Vector3 &pos, const Material &m) : SceneObject(m), m_radius(r) { position_ = pos; }
  double GetRadius() const { return m_radius; }
  // Sphere::CheckIntersection (src/renderer/Sphere.h:21-78), uv when textured
  bool CheckIntersection(const Ray &ray, HitInformation &hit) const override;

 protected:
  double m_radius;
};

// LightBase (src/renderer/LightBase.h): light sampling for next-event
// estimation.
class LightBase {
 public:
  virtual ~LightBase() {}
  virtual void SampleOnePoint(Vector3 &point, Vector3 &normal, double &pdf, const Random &rnd) const = 0;
  virtual bool SampleOnePointWithTargetPoint(Vector3 &sampledPoint, Vector3 &sampledPointNormal, double &pdf,
                                             const Vector3 &targetPoint, const Random &rnd) const = 0;
  virtual double TotalPower() const = 0;
};
// SphereLight (src/renderer/SphereLight.h:8-79): an emissive sphere the
// renderer sees as a light.
class SphereLight : public Sphere, public LightBase {
 public:
  SphereLight(double r, const Vector3 &pos, const Material &m) : Sphere(r, pos, m) {}
  // uniform point on the sphere (:17-35)
  void SampleOnePoint(Vector3 &point, Vector3 &normal, double &pdf, const Random &rnd) const override;
  // uniform over the cone of the sphere seen from targetPoint (:38-66)
  bool SampleOnePointWithTargetPoint(Vector3 &sampledPoint, Vector3 &sampledPointNormal, double &pdf,
                                     const Vector3 &targetPoint, const Random &rnd) const override;
  // 4 pi r^2 x the rate-weighted emission length (:68-75)
  double TotalPower() const override;
};

class Polygon : public SceneObject {
 public:
  Polygon(const Vector3 &pos1, const Vector3 &pos2, const Vector3 &pos3, const Vector3 &uv1, const Vector3 &uv2,
          const Vector3 &uv3, const Vector3 &normal1, const Vector3 &normal2, const Vector3 &normal3,
          const Material &mat, const Vector3 &pos);
  // Polygon::Transform (src/renderer/Polygon.h:51-61): rotate vertex and
  // edges, scale per axis, translate through position_; normals rotate only.
  void Transform(const Vector3 &pos, const Vector3 &scale = Vector3::One(), const Matrix &rot = Matrix::Identity());
  const Vector3 *PosAndEdges() const { return m_posAndEdges; }
  const Vector3 *UVOrigAndEdges() const { return m_uvOrigAndEdges; }
  const Vector3 *NormalAndDiffs() const { return m_normalAndDiffs; }
  // Polygon::CalculateNormal (src/renderer/Polygon.h:45-49)
  static Vector3 CalculateNormal(const Vector3 &v0, const Vector3 &v1, const Vector3 &v2);
  // Polygon::CheckIntersection (src/renderer/Polygon.h:63-106): single-sided
  bool CheckIntersection(const Ray &ray, HitInformation &hit) const override;

 private:
  Vector3 m_posAndEdges[3];
  Vector3 m_uvOrigAndEdges[3];
  Vector3 m_normalAndDiffs[3];
};

// Aperture (src/renderer/Aperture.h:9-13).  The GPU path renders circular
// apertures: Radius() is the circle's radius, or negative for an aperture it
// cannot render (PathTracer then reports an error).
class Aperture {
 public:
  virtual ~Aperture() {}
  virtual void SampleOnePoint(double &sampledX, double &sampledY, const Random &rnd) const = 0;
  virtual double Radius() const { return -1.0; }
};
// CircleAperture (src/renderer/Aperture.h:16-37): uniform on the disc.
class CircleAperture : public Aperture {
 public:
  explicit CircleAperture(double radius = 0.0) : m_radius(radius) {}
  void SampleOnePoint(double &sampledX, double &sampledY, const Random &rnd) const override {
    const double theta = 2 * PI * rnd.nextDouble();
    const double r = m_radius * std::sqrt(rnd.nextDouble());
    sampledX = r * std::cos(theta);
    sampledY = r * std::sin(theta);
  }
  double Radius() const override { return m_radius; }

 private:
  double m_radius;
};

class Camera {
 public:
  Camera(size_t screenWidth, size_t screenHeight, const Vector3 &cameraPos = Vector3(50.0, 52.0, 220.0),
         const Vector3 &cameraDir = Vector3(0.0, -0.04, -1.0), const Vector3 &cameraUp = Vector3(0.0, 1.0, 0.0),
         double screenHeightIn3Dcoordinate = 30.0, double distanceToScreen = 40, double focalDistance = 200);
  void SetAperture(std::shared_ptr<Aperture> a) { m_aperture = a; }
  size_t GetScreenWidth() const { return m_width; }
  size_t GetScreenHeight() const { return m_height; }
  const Vector3 &GetCameraPosition() const { return m_camPos; }
  const Vector3 &GetScreenCenterPosition() const { return m_screenCenter; }
  const Vector3 &GetScreenAxisXvector() const { return m_screenXaxis; }
  const Vector3 &GetScreenAxisYvector() const { return m_screenYaxis; }
  // Camera::SampleRayForPixel (src/renderer/Camera.h:65-93): the ray through
  // screen point (x, y), through a lens point when an aperture is set
  Ray SampleRayForPixel(double x, double y, const Random &rnd) const;
  // the device form; aperture_radius < 0 when the aperture is not circular
  omr_camera ToDevice() const;

 private:
  size_t m_width, m_height;
  Vector3 m_camPos, m_camDir, m_camUp;
  double m_nearScreenHeight, m_distToScreen, m_focalDistance;
  Vector3 m_screenXaxis, m_screenYaxis, m_screenCenter;
  std::shared_ptr<Aperture> m_aperture;
};

// Infinite horizontal plane y = const, two-sided (InfiniteFloor,
// src/renderer/InfiniteFloor.h:10-45); never inside the BVH.
class InfiniteFloor : public SceneObject {
 public:
  InfiniteFloor(double y, const Material &m) : SceneObject(m), m_y(y) {}
  double GetY() const { return m_y; }
  bool CheckIntersection(const Ray &ray, HitInformation &hit) const override;

 private:
  double m_y;
};

// Radiance RGBE image (HDRImage, src/tools/HDRImage.h:8-81): read and write,
// run-length encoded scanlines as the reference writes them.
class HDRImage : public Image {
 public:
  bool ReadFromRadianceFile(const std::string &file);
  bool WriteToRadianceFile(const std::string &file);
  // replaces the pixels (row-major, row 0 first)
  void SetImage(size_t width, size_t height, std::vector<Color> pixels);
  static void Color2RGBE(unsigned char rgbe[4], const Color &c);
  static void RGBE2Color(const unsigned char rgbe[4], Color &c);

 private:
  struct Header {
    int valid = 0;  // 1 program type, 2 gamma, 4 exposure
    std::string programtype;
    float gamma = 0.0f, exposure = 0.0f;
    int height = 0, width = 0;
  };
  Header m_info;
};

// Latlong environment (IBL, src/renderer/IBL.h:14-45).
class IBL {
 public:
  IBL(uint32_t w, uint32_t h, std::vector<float> rgb) : m_w(w), m_h(h), m_rgb(std::move(rgb)) {}
  // IBL(hdr_filename) (IBL.h:16-26): reads a Radiance file; empty when it fails
  explicit IBL(const std::string &hdr_filename);
  bool IsValid() const { return m_w > 0 && m_h > 0; }
  uint32_t Width() const { return m_w; }
  uint32_t Height() const { return m_h; }
  const std::vector<float> &Pixels() const { return m_rgb; }

 private:
  uint32_t m_w, m_h;
  std::vector<float> m_rgb;
};

class Model;

// A scene flattened into what omr_scene_upload takes: triangles, then
// spheres, then infinite floors; `desc` points into the vectors.
struct SceneArrays {
  std::vector<float> tri, nrm, sph, planes, uv;
  std::vector<uint32_t> prim_mat, prim_mat_count;
  std::vector<omr_material> mats;
  std::vector<SceneObject *> prims;        // primitive id -> object
  std::vector<std::vector<float>> tex_rgb; // textured images as float RGB
  std::vector<omr_texture> textures;
  omr_scene_desc desc;
};

class Scene {
 public:
  // Scene::IntersectionInformation (src/scenes/Scene.h:20-24)
  struct IntersectionInformation {
    HitInformation hit;
    SceneObject *object = nullptr;
    Color texturedHitpointColor;
  };
  virtual ~Scene();
  void ConstructBVH() { ConstructQBVH(); }
  // Builds the wide BVH and keeps the scene resident on the GPU.
  void ConstructQBVH();
  // Scene::CheckIntersection (src/scenes/Scene.cpp:51-85): the closest object
  // (found by the GPU traversal), its hit record computed by the object in
  // FP64 as the reference does.  Slow per ray; rendering does not use it.
  bool CheckIntersection(const Ray &ray, IntersectionInformation &info) const;
  // Scene::GetLights (src/scenes/Scene.h:36): the objects added that are lights
  const std::vector<LightBase *> GetLights() const { return m_lights; }
  const IBL *GetIBL() const { return m_ibl.get(); }
  virtual Color Background() const { return Color(0, 0, 0); }
  virtual bool IsValid() const { return true; }

  // Device handle for the current contents (uploads on first use).
  omr_scene_id Upload(omr_ctx *ctx) const;
  // The current contents as upload arrays (no device needed).
  void Describe(SceneArrays &out) const;

 protected:
  Scene() {}
  void AddObject(SceneObject *obj, bool doDelete = true, bool containedInBVH = true);
  // Scene::AddModel (src/scenes/Scene.h:78-88): every polygon of the model,
  // material by material; the scene keeps the model alive when doDelete.
  void AddModel(Model *model, bool doDelete = true, bool containedInBVH = true);
  // Scene::AddInfiniteFLoor (src/scenes/Scene.h:71)
  SceneObject *AddInfiniteFLoor(double y, const Material &material);
  std::vector<SceneObject *> AddFloorXZ_yUp(double size_x, double size_z, const Vector3 &position, const Material &m);
  std::vector<SceneObject *> AddFloorXY_zUp(double size_x, double size_y, const Vector3 &position, const Material &m);
  std::vector<SceneObject *> AddFloorYZ_xUp(double size_y, double size_z, const Vector3 &position, const Material &m);
  std::vector<SceneObject *> AddFloorXZ_yDown(double size_x, double size_z, const Vector3 &position, const Material &m);
  std::vector<SceneObject *> AddFloorXY_zDown(double size_x, double size_y, const Vector3 &position, const Material &m);
  std::vector<SceneObject *> AddFloorYZ_xDown(double size_y, double size_z, const Vector3 &position, const Material &m);
  void SetIBL(std::unique_ptr<IBL> ibl) { m_ibl = std::move(ibl); }

 private:
  struct Owned {
    SceneObject *obj;
    bool doDelete;
  };
  std::vector<Owned> m_objects;
  std::vector<LightBase *> m_lights;
  std::vector<std::unique_ptr<Model>> m_models;
  std::unique_ptr<IBL> m_ibl;
  mutable omr_ctx *m_ctx = nullptr;
  mutable omr_scene_id m_id = -1;
  mutable std::vector<SceneObject *> m_prims;  // primitive id -> object
};

// Wavefront OBJ mesh with MTL materials (Model, src/renderer/Model.cpp:50-367):
// faces of n >= 3 vertices as a fan, v, v/vt, v//vn, v/vt/vn indices (1-based),
// per-face normals when a face gives none, one polygon list per material.
class Model {
 public:
  typedef Polygon *PolygonPtr;
  typedef std::vector<PolygonPtr> PolygonList;
  Model() {}
  ~Model();
  bool ReadFromObj(const std::string &filename, bool flipV_of_UV = false);
  void Transform(const Vector3 &pos, const Vector3 &scale = Vector3::One(), const Matrix &rot = Matrix::Identity());
  const Vector3 &GetPosition() const { return m_position; }
  size_t GetMaterialCount() const { return m_materials.size(); }
  const Material &GetMaterial(size_t i) const { return m_materials.at(i); }
  const PolygonList &GetPolygonList(const Material &mat) const;

 private:
  void Clear();
  std::vector<Material> m_materials;   // one per material name, in the reference's listing order
  std::vector<Material> m_list_keys;   // distinct material values that own a polygon list
  std::vector<PolygonList> m_meshes;   // parallel to m_list_keys
  Vector3 m_position;
};

// Scene read from the reference's text format (SceneFromExternalFile,
// src/scenes/SceneFromExternalFile.cpp:13-531): header (Base Dir, IBL, Space
// Partitioning) up to "Objects:", then blocks "Obj Mesh", "Sphere", "Floor",
// "SphereLight", "InfiniteFloor" of "Key: value" lines.  IsValid() is false
// when the file or any block fails to parse (messages on stderr).
class SceneFromExternalFile : public Scene {
 public:
  explicit SceneFromExternalFile(const std::string &file);
  bool IsValid() const override { return m_isValid; }
  const std::string &GetIBLFileName() const { return m_iblFileName; }

 private:
  typedef std::pair<std::string, std::string> LinePair;
  bool ReadFromFile(const std::string &file);
  bool ReadHeader(const std::vector<LinePair> &lines);
  bool ParseMaterial(const std::string &data, Material &mat);
  bool ProcessMaterials(SceneObject *obj, std::vector<Material> &materials, std::vector<double> &rates);
  bool ReadMesh(const std::vector<LinePair> &lines);
  bool ReadSphere(const std::vector<LinePair> &lines, bool isLight);
  bool ReadFloor(const std::vector<LinePair> &lines);
  bool ReadInfiniteFloor(const std::vector<LinePair> &lines);
  bool m_isValid = false;
  std::string m_baseDir, m_iblFileName, m_spacePartitioningMethod;
};

// ScreenPixels (src/renderer/ScreenPixels.hpp:18-78, ScreenPixels.cpp:37-130):
// xCount x yCount viewports of w x h pixels in one buffer; viewport (x, y)
// coordinates map to buffer index X + (Height - Y - 1) x Width with X, Y the
// screen coordinates, so rows are stored bottom-up.  PathTracer fills
// viewport (0, 0) of a 2 x 1 layout, as the reference does.
class ScreenPixels {
 public:
  class ScreenViewport {
   public:
    void SetColor(size_t x, size_t y, const Color &color) { m_screen->m_pixels[calcIndex(x, y)] = color; }
    const Color &GetColor(size_t x, size_t y) const { return m_screen->m_pixels[calcIndex(x, y)]; }

   private:
    friend class ScreenPixels;
    size_t calcIndex(size_t x, size_t y) const {
      const size_t W = m_screen->CalcScreenWidth(), H = m_screen->CalcScreenHeight();
      return (x + m_xOffset) + (H - (y + m_yOffset) - 1) * W;
    }
    size_t m_xOffset = 0, m_yOffset = 0;
    ScreenPixels *m_screen = nullptr;
  };
  void InitializePixels(size_t widthPerViewport, size_t heightPerViewport, size_t xViewportCount,
                        size_t yViewportCount) {
    m_w = widthPerViewport;
    m_h = heightPerViewport;
    m_xCount = xViewportCount;
    m_yCount = yViewportCount;
    m_pixels.assign(m_w * m_h * m_xCount * m_yCount, Color());
    m_views.assign(m_xCount * m_yCount, ScreenViewport());
    for (size_t y = 0; y < m_yCount; y++)
      for (size_t x = 0; x < m_xCount; x++) {
        ScreenViewport &v = m_views[x + y * m_xCount];
        v.m_screen = this;
        v.m_xOffset = x * m_w;
        v.m_yOffset = y * m_h;
      }
  }
  ScreenViewport *GetViewport(size_t xViewport, size_t yViewport) {
    if (xViewport >= m_xCount || yViewport >= m_yCount) return nullptr;
    return &m_views[xViewport + yViewport * m_xCount];
  }
  size_t CalcScreenWidth() const { return m_xCount * m_w; }
  size_t CalcScreenHeight() const { return m_yCount * m_h; }
  const Color *GetScreen() const { return m_pixels.data(); }
  Color *GetScreen() { return m_pixels.data(); }

 private:
  size_t m_w = 0, m_h = 0, m_xCount = 0, m_yCount = 0;
  std::vector<Color> m_pixels;
  std::vector<ScreenViewport> m_views;
};

class Renderer {
 protected:
  std::atomic<bool> m_enableRendering{true};

 public:
  virtual ~Renderer() {}
  virtual void RenderScene(const Scene &scene) = 0;
  virtual const ScreenPixels *GetResult() const = 0;
  virtual const int GetCurrentSampleCount() const = 0;
  virtual std::string GetCurrentRenderingInfo() const { return ""; }
  void StopRendering() { m_enableRendering = false; }
};

// PathTracer (src/renderer/PathTracer.h:18-88) on the GPU.
class PathTracer : public Renderer {
 public:
  typedef std::function<void(int samples, const Color *result, size_t width, size_t height,
                             double accumulatedRenderingTime)>
      RenderingFinishCallbackFunction;
  PathTracer(const Camera &camera, int samples, int supersamples);
  PathTracer(const Camera &camera, int min_samples, int max_samples, int step, int supersamples,
             RenderingFinishCallbackFunction callback);
  virtual ~PathTracer();
  // PathTracer::EnableNextEventEstimation (PathTracer.h:30-33), with the
  // reference's semantics (PathTracer.cpp:406-439): surface emission counts
  // only where the camera ray hits, and the direct-light term is zero.
  void EnableNextEventEstimation(bool enable = true) { m_nee = enable; }
  // Environment importance sampling with shadow rays and MIS (no reference
  // counterpart: the reference's IBL importance map is never built).
  void EnableEnvironmentImportanceSampling(bool enable = true) { m_envIS = enable; }
  void SetMaxThreadCount(size_t) {}  // threads are GPU lanes
  // Depth parameters (reference constants MinDepth = 5, MaxDepth = 64).
  void SetDepth(uint32_t rr_depth, uint32_t max_depth) { m_rrDepth = rr_depth; m_maxDepth = max_depth; }
  void SetSeed(uint64_t seed) { m_seed = seed; }
  void RenderScene(const Scene &scene) override;
  const ScreenPixels *GetResult() const override { return m_result.get(); }
  const int GetCurrentSampleCount() const override { return m_currentSamples; }
  std::string GetCurrentRenderingInfo() const override;
  // Last error from the library (empty when fine).
  const std::string &LastError() const { return m_error; }

 private:
  Camera m_camera;
  int m_min, m_max, m_step, m_ss;
  int m_currentSamples = 0, m_previousSamples = 0;
  RenderingFinishCallbackFunction m_callback;
  std::unique_ptr<ScreenPixels> m_result;
  bool m_nee = false, m_envIS = false;
  uint32_t m_rrDepth = 5, m_maxDepth = 0;
  uint64_t m_seed = 0x6F6D7233ull;
  uint64_t m_segments = 0;
  omr_ctx *m_ctx = nullptr;
  std::string m_error;
};

// SceneFactory / SceneFactoryManager (src/scenes/SceneFactory.h:7-48)
class SceneFactory {
 public:
  virtual ~SceneFactory() {}
  virtual std::shared_ptr<Scene> Create(const std::string &info) const = 0;
};
class SceneFactoryManager {
 public:
  static SceneFactoryManager &GetInstance() {
    static SceneFactoryManager m;
    return m;
  }
  void Register(const std::string &name, std::shared_ptr<SceneFactory> f) { m_factories[name] = f; }
  std::shared_ptr<SceneFactory> Get(const std::string &name) const {
    auto it = m_factories.find(name);
    return it == m_factories.end() ? nullptr : it->second;
  }

 private:
  std::map<std::string, std::shared_ptr<SceneFactory>> m_factories;
};

// Process-wide library context used by the drop-in classes (device 0).
omr_ctx *DefaultContext();

}  // namespace OmochiRenderer
