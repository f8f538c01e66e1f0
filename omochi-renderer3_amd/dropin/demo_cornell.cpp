// demo_cornell.cpp -- the reference's own Cornell box scene
// (src/scenes/CornellBoxScene.cpp:10-36) through the drop-in classes, rendered
// progressively as src/main.cpp does.  Prints one line per sample batch and a
// final JSON summary; exits non-zero on any error.
#include <cmath>
#include <cstdio>

#include "OmochiRenderer.h"

using namespace OmochiRenderer;

namespace {

class CornellBoxScene : public Scene {
 public:
  CornellBoxScene() {
    typedef Material M;
    AddFloorYZ_xUp(200, 200, Vector3(1, 40.8, 81.6), M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.75, 0.25, 0.25)));
    AddFloorYZ_xDown(200, 200, Vector3(99, 40.8, 81.6),
                     M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.25, 0.25, 0.75)));
    AddFloorXY_zUp(150, 150 / 1.77777777777, Vector3(50, 40.8, 0),
                   M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(1, 1, 1)));
    AddFloorXY_zDown(150, 150 / 1.77777777777, Vector3(50, 40.8, 250),
                     M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.75, 0.75, 0.75)));
    AddFloorXZ_yUp(2000, 2000, Vector3(50, 0, 81.6), M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.75, 0.75, 0.75)));
    AddFloorXZ_yDown(200, 200, Vector3(50, 81.6, 81.6),
                     M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.75, 0.75, 0.75)));
    SceneObject *obj = new Sphere(20, Vector3(50, 20, 50), M(M::REFLECTION_TYPE_LAMBERT, Color(), Color(0.25, 0.75, 0.25)));
    obj->AddMaterial(M(M::REFLECTION_TYPE_SPECULAR, Color(), Color(0.99, 0.99, 0.99)), 0.1);
    AddObject(obj);
    AddObject(new Sphere(16.5, Vector3(19, 16.5, 25), M(M::REFLECTION_TYPE_SPECULAR, Color(), Color(0.99, 0.99, 0.99))));
    AddObject(new Sphere(16.5, Vector3(77, 16.5, 78),
                         M(M::REFLECTION_TYPE_REFRACTION, Color(), Color(0.99, 0.99, 0.99), 1.5)));
    AddObject(new SphereLight(15, Vector3(50.0, 90, 81.6), M(M::REFLECTION_TYPE_LAMBERT, Color(36, 36, 36), Color())));
    ConstructQBVH();
  }
};

}  // namespace

int main(int argc, char **argv) {
  int w = argc > 1 ? atoi(argv[1]) : 64, h = argc > 2 ? atoi(argv[2]) : 36, batches = argc > 3 ? atoi(argv[3]) : 3;
  // settings.txt: camera position 50,55,230 direction 0.00001,-0.15,-1.0, supersamples 4
  Camera camera(w, h, Vector3(50, 55, 230), Vector3(0.00001, -0.15, -1.0));
  int calls = 0;
  double last_mean = 0;
  PathTracer pt(camera, 1, batches, 1, 2,
                [&](int samples, const Color *img, size_t w2, size_t hh, double) {
                  double s = 0;
                  for (size_t y = 0; y < hh; y++)
                    for (size_t x = 0; x < w2 / 2; x++) s += img[x + y * w2].x + img[x + y * w2].y + img[x + y * w2].z;
                  last_mean = s / (3.0 * (w2 / 2) * hh);
                  calls++;
                  std::printf("samples %d mean %.6f\n", samples, last_mean);
                });
  CornellBoxScene scene;
  pt.RenderScene(scene);
  if (!pt.LastError().empty()) {
    std::fprintf(stderr, "render failed: %s\n", pt.LastError().c_str());
    return 1;
  }
  Scene::IntersectionInformation info;
  bool hit = scene.CheckIntersection(Ray(Vector3(50, 40, 200), Vector3(0, 0, -1)), info);
  std::printf("{\"batches\": %d, \"mean\": %.6f, \"probe_hit\": %d, \"probe_t\": %.4f}\n", calls, last_mean,
              hit ? 1 : 0, info.hit.distance);
  return (calls == batches && std::isfinite(last_mean) && last_mean > 0 && hit) ? 0 : 2;
}
