// A caller of the reference, built against the C++ drop-in: the render path
// of the reference's src/main.cpp:98-136 (Camera + CircleAperture, PathTracer
// with a finish callback, SceneFactoryManager) over the reference's own
// CornellBoxScene.h / .cpp, whose #include lines are the only change (the
// recipe is oracle/Makefile's dropin_compat target), and probes of the
// interfaces such a caller uses: Scene::IntersectionInformation, GetLights,
// Camera::SampleRayForPixel, ScreenPixels viewports, Random and the names of
// src/tools/Constant.h.
//
//   dropin_compat render            renders 64x36 at 2 samples, prints
//                                   "render <callbacks> <width> <height>"
//   dropin_compat <rays> <hits>     prints the probes, then for every ray
//                                   (8 floats: origin, pad, direction, pad)
//                                   writes 9 doubles: distance (-1 = miss),
//                                   normal, material count, and the first
//                                   material's type, color.x, emission.x, rate
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "OmochiRenderer.h"
#include "CornellBoxScene.h"  // the reference's own header, #include lines removed

using namespace OmochiRenderer;

int main(int argc, char **argv) {
  // src/main.cpp:98-136, with the settings.txt values for the Cornell box
  Camera camera(64, 36, Vector3(50, 55.0, 230.0), Vector3(0.00001, -0.15, -1.0), Vector3(0, 1, 0), 30.0, 40.0, 300);
  camera.SetAperture(std::shared_ptr<Aperture>(new CircleAperture(1.3)));
  int calls = 0;
  PathTracer::RenderingFinishCallbackFunction callback(
      [&calls](int samples, const Color *img, size_t width, size_t height, double accumulatedRenderingTime) {
        calls++;
        (void)samples; (void)img; (void)width; (void)height; (void)accumulatedRenderingTime;
      });
  std::shared_ptr<PathTracer> renderer = std::make_shared<PathTracer>(camera, 1, 2, 1, 2, callback);
  renderer->EnableNextEventEstimation(false);
  renderer->SetMaxThreadCount(6);
  SceneFactoryManager::GetInstance().Register("CornellBoxScene", std::make_shared<CornellBoxSceneFactory>());
  auto factory = SceneFactoryManager::GetInstance().Get("CornellBoxScene");
  if (!factory) return 3;
  std::shared_ptr<Scene> scene = factory->Create("");
  if (argc > 1 && std::strcmp(argv[1], "render") == 0) {
    renderer->RenderScene(*scene);
    const ScreenPixels *px = renderer->GetResult();
    std::printf("render %d %zu %zu\n", calls, px ? px->CalcScreenWidth() : 0, px ? px->CalcScreenHeight() : 0);
    return renderer->LastError().empty() && calls == 2 ? 0 : 4;
  }
  // interfaces a caller of the reference uses
  Random rnd(12345);
  Ray r = camera.SampleRayForPixel(10.5, 3.5, rnd);
  std::printf("lights %zu ray %.17g %.17g %.17g const %.17g %.17g %.17g %.17g %.17g\n", scene->GetLights().size(),
              r.dir.x, r.dir.y, r.dir.z, (double)INF, (double)PI, (double)EPS, (double)REFRACTIVE_INDEX_VACUUM,
              (double)REFRACTIVE_INDEX_OBJECT);
  ScreenPixels sp;
  sp.InitializePixels(4, 3, 2, 1);
  sp.GetViewport(1, 0)->SetColor(1, 0, Color(7, 8, 9));
  const Color *s = sp.GetScreen();
  std::printf("viewport %g %zu %zu\n", s[(4 + 1) + (3 - 0 - 1) * 8].x, sp.CalcScreenWidth(), sp.CalcScreenHeight());
  // hits: rays from argv[1] (n x 8 floats); t, normal, first material per ray
  FILE *f = std::fopen(argv[1], "rb");
  FILE *o = std::fopen(argv[2], "wb");
  float ray[8];
  while (std::fread(ray, sizeof(float), 8, f) == 8) {
    Scene::IntersectionInformation info;
    bool hit = scene->CheckIntersection(Ray(Vector3(ray[0], ray[1], ray[2]), Vector3(ray[4], ray[5], ray[6])), info);
    double rec[9] = {hit ? info.hit.distance : -1.0, info.hit.normal.x, info.hit.normal.y, info.hit.normal.z,
                     hit ? (double)info.object->GetMaterialCount() : 0.0, 0, 0, 0, 0};
    if (hit) {
      const Material &m = info.object->GetMaterial(0)->material_;
      rec[5] = m.reflection_type;
      rec[6] = m.color.x;
      rec[7] = m.emission.x;
      rec[8] = info.object->GetMaterial(0)->rate_;
    }
    std::fwrite(rec, sizeof(double), 9, o);
  }
  std::fclose(f);
  std::fclose(o);
  return 0;
}
