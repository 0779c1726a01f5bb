// buffers.hpp -- output and input buffers of the host scene API
// (host_includes/buffers.hpp:13-138).  The two output buffers map to the
// library's FP64 accumulation and its display image.
#pragma once

#include "host_common.hpp"

// buffers.hpp:13-18
inline Buffer createFrameBuffer(int Nx, int Ny, Context& g_context) {
  Buffer b = g_context->createBuffer(RT_BUFFER_OUTPUT);
  b->setFormat(RT_FORMAT_FLOAT4);
  b->setSize(Nx, Ny);
  return b;
}
// buffers.hpp:21-26
inline Buffer createDisplayBuffer(int Nx, int Ny, Context& g_context) {
  Buffer b = g_context->createBuffer(RT_BUFFER_OUTPUT);
  b->setFormat(RT_FORMAT_UNSIGNED_BYTE4);
  b->setSize(Nx, Ny);
  return b;
}
// buffers.hpp:33-138: typed input buffers hold host copies
template <class T>
inline Buffer createBuffer(const std::vector<T>& list, Context& g_context) {
  Buffer b = g_context->createBuffer(RT_BUFFER_INPUT);
  b->setSize(list.size());
  b->host.resize(list.size() * sizeof(T));
  if (!list.empty()) std::memcpy(b->host.data(), list.data(), b->host.size());
  return b;
}
