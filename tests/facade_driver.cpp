// Drives the host scene API the way the reference's main.cpp does
// (main.cpp:26-39 renderFrame, :56-63 Optix_Config, :264-271 the frame loop
// and displayBuffer->map()): a scene function records the scene through the
// OptiX-shaped facade, the output buffers are created and bound, an empty
// launch builds the acceleration structures, then every frame sets
// context["frame"] and launches W x H.  Writes accBuffer->map() (float4 per
// pixel: summed RGB and the frame count, rows top to bottom) and
// displayBuffer->map() (uchar4) to a file.  Built by tests/Makefile; used
// by tests/test_gpu_launch.py.
//
//   facade_driver <Cornell|RandomSpheres> W H frames out.bin
#include <cstdio>
#include <cstdlib>
#include <string>

#include "buffers.hpp"
#include "config_scenes.hpp"
#include "image_save.hpp"

static float renderFrame(Context& ctx, int Nx, int Ny) {
  ctx->validate();
  ctx->launch(0, Nx, Ny);
  return 0.f;
}

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  const std::string which = argv[1];
  App_State app;
  app.W = std::atoi(argv[2]);
  app.H = std::atoi(argv[3]);
  const int frames = std::atoi(argv[4]);
  try {
    SceneArena A;
    ConfigDefaults d{};
    if (which == "Cornell") d = CornellBox(app, A);
    else if (which == "RandomSpheres") d = RandomSpheres(app, A);
    else return 2;
    app.context->setRayTypeCount(2);
    app.context->setMaxTraceDepth(5);
    app.context["samples"]->setInt(frames);
    app.context["russian"]->setInt(d.russian ? 1 : 0);
    app.context["maxDepth"]->setInt(d.depth);
    app.context["rr_min_depth"]->setInt(d.rr_min_depth);
    app.accBuffer = createFrameBuffer(app.W, app.H, app.context);
    app.context["acc_buffer"]->set(app.accBuffer);
    app.displayBuffer = createDisplayBuffer(app.W, app.H, app.context);
    app.context["display_buffer"]->set(app.displayBuffer);
    renderFrame(app.context, 0, 0);  // the build launch
    for (int f = 0; f < frames; f++) {
      app.context["frame"]->setInt(f);
      renderFrame(app.context, app.W, app.H);
    }
    const size_t n = (size_t)app.W * app.H;
    FILE* fp = std::fopen(argv[5], "wb");
    if (!fp) return 3;
    const void* acc = app.accBuffer->map();
    std::fwrite(acc, sizeof(float) * 4, n, fp);
    app.accBuffer->unmap();
    const void* disp = app.displayBuffer->map();
    std::fwrite(disp, 4, n, fp);
    app.displayBuffer->unmap();
    std::fclose(fp);
  } catch (const char* e) {
    std::fprintf(stderr, "error: %s\n", e);
    return 1;
  } catch (const std::string& e) {
    std::fprintf(stderr, "error: %s\n", e.c_str());
    return 1;
  }
  return 0;
}
