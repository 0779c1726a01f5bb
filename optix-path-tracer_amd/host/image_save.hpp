// image_save.hpp -- file output of the host scene API (host_includes/image_save.hpp:7-73).
// Save_PNG / Save_HDR keep the reference's signatures (App_State and the
// accumulation buffer; the extension is appended to app.fileName) and hand
// the work to pt_save_image on the context the buffer belongs to: the
// average of the FP64 accumulation over the frames rendered so far (the
// reference divides by app.samples, the same number once the render loop
// of main.cpp:264-271 has finished), sqrt gamma and 255.99 scaling for PNG,
// c / (1 + c) for Radiance HDR, rows top first.  Return: 1 on success, 0
// when the file cannot be written or nothing was rendered.
#pragma once

#include <string>

#include "host_common.hpp"

inline int Save_PNG(App_State& app, Buffer& buffer) {
  app.fileName += ".png";
  return buffer->ctx->save_image(app.fileName.c_str(), PT_IMAGE_PNG) ? 1 : 0;
}

inline int Save_HDR(App_State& app, Buffer& buffer) {
  app.fileName += ".hdr";
  return buffer->ctx->save_image(app.fileName.c_str(), PT_IMAGE_HDR) ? 1 : 0;
}
