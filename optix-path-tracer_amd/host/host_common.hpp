// host_common.hpp -- App_State, scene RNG and Light_Sampler of the host scene
// API (host_includes/host_common.hpp:20-82), on the C-ABI facade.
#pragma once

#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "optix_facade.hpp"

// host_common.hpp:21-45 (defaults kept: 500x500, 500 spp, scene 2, RR on,
// depth 50)
struct App_State {
  App_State() {
    context = Context::create();
    W = H = 500;
    samples = 500;
    scene = 2;
    model = 0;
    frequency = 1;
    currentSample = 0;
    showProgress = true;
    RTX = true;
    russian = true;
    depth = 50;
    start = done = false;
    fileType = 0;
    fileName = "out";
  }
  Context context;
  int W, H, samples, scene, currentSample, model, frequency, fileType, depth;
  bool done, start, showProgress, RTX, russian;
  Buffer accBuffer, displayBuffer;
  std::string fileName;
};

// host_common.hpp:59-63.  The reference draws std::mt19937(0) through
// uniform_real_distribution<float>, whose mapping is implementation defined;
// here each draw is the top 24 bits of one mt19937 output times 2^-24.
inline std::mt19937& rnd_engine() {
  static std::mt19937 gen(0);
  return gen;
}
inline void rnd_reseed(uint32_t seed) { rnd_engine().seed(seed); }
inline float rnd() { return (float)(rnd_engine()() >> 8) * (1.0f / 16777216.0f); }

// host_common.hpp:65-68
struct Light_Sampler {
  std::vector<Program> sample, pdf;
  std::vector<float3> emissions;
};

// host_common.hpp:71-82
inline float saturate(float x) { return ffmax(0.f, ffmin(1.f, x)); }
inline float Sign(float x) { return x < 0.0f ? -1.0f : (x > 0.0f ? 1.0f : 0.0f); }

// host_common.hpp:48-57: programs are named device entry points in the
// reference; here they are opaque values recorded by name.
inline Program createProgram(const char* /*file*/, const std::string& name, Context& /*ctx*/) {
  auto p = std::make_shared<ptfacade::ProgramObj>();
  p->name = name;
  return Program(p);
}
