// obj_loader.hpp -- Wavefront OBJ/MTL reader for the host scene API: the
// subset mesh.hpp:57-149 takes from tinyobjloader (v, vn, vt, f with
// polygon fan triangulation, negative indices, usemtl, mtllib with Ka/Kd).
// Faces are flattened into unindexed triangles (i, i+1, i+2).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

struct ObjMaterial {
  std::string name;
  float ambient[3] = {0, 0, 0};
  float diffuse[3] = {0, 0, 0};
  std::string diffuse_texname;
};

struct ObjMesh {
  std::vector<float> v, n, uv;   // flattened per corner
  std::vector<uint32_t> idx;     // (i, i+1, i+2) per triangle
  std::vector<int> face_material;
  std::vector<ObjMaterial> materials;
};

inline bool obj_load_mtl(const std::string& path, std::vector<ObjMaterial>& mats, std::map<std::string, int>& by_name) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "newmtl") {
      ObjMaterial m;
      ss >> m.name;
      by_name[m.name] = (int)mats.size();
      mats.push_back(m);
    } else if (!mats.empty() && tag == "Ka") {
      ss >> mats.back().ambient[0] >> mats.back().ambient[1] >> mats.back().ambient[2];
    } else if (!mats.empty() && tag == "Kd") {
      ss >> mats.back().diffuse[0] >> mats.back().diffuse[1] >> mats.back().diffuse[2];
    } else if (!mats.empty() && tag == "map_Kd") {
      ss >> mats.back().diffuse_texname;
    }
  }
  return true;
}

inline bool obj_load(const std::string& path, const std::string& folder, ObjMesh& out, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  std::vector<float> P, N, T;
  std::map<std::string, int> mat_by_name;
  int cur_mat = -1;
  bool any_n = false, any_t = false, all_n = true, all_t = true;
  struct Corner { long v, t, n; };
  std::vector<std::vector<Corner>> tris;
  std::vector<int> tri_mat;
  std::string line;
  size_t lineno = 0;
  auto fix = [](long i, size_t count) -> long { return i > 0 ? i - 1 : (i < 0 ? (long)count + i : -1); };
  while (std::getline(in, line)) {
    lineno++;
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "v") {
      float x, y, z;
      ss >> x >> y >> z;
      P.insert(P.end(), {x, y, z});
    } else if (tag == "vn") {
      float x, y, z;
      ss >> x >> y >> z;
      N.insert(N.end(), {x, y, z});
    } else if (tag == "vt") {
      float u = 0, v = 0;
      ss >> u >> v;
      T.insert(T.end(), {u, v});
    } else if (tag == "f") {
      std::vector<Corner> poly;
      std::string tok;
      while (ss >> tok) {
        Corner c{-1, -1, -1};
        long a = 0, b = 0, d = 0;
        if (std::sscanf(tok.c_str(), "%ld/%ld/%ld", &a, &b, &d) == 3) { c = {a, b, d}; }
        else if (std::sscanf(tok.c_str(), "%ld//%ld", &a, &d) == 2) { c = {a, 0, d}; }
        else if (std::sscanf(tok.c_str(), "%ld/%ld", &a, &b) == 2) { c = {a, b, 0}; }
        else if (std::sscanf(tok.c_str(), "%ld", &a) == 1) { c = {a, 0, 0}; }
        else { err = "bad face token at line " + std::to_string(lineno); return false; }
        c.v = fix(c.v, P.size() / 3);
        c.t = fix(c.t, T.size() / 2);
        c.n = fix(c.n, N.size() / 3);
        if (c.v < 0 || (size_t)c.v >= P.size() / 3) { err = "vertex index out of range at line " + std::to_string(lineno); return false; }
        poly.push_back(c);
      }
      for (size_t k = 2; k < poly.size(); k++) {
        tris.push_back({poly[0], poly[k - 1], poly[k]});
        tri_mat.push_back(cur_mat);
      }
    } else if (tag == "usemtl") {
      std::string name;
      ss >> name;
      auto it = mat_by_name.find(name);
      cur_mat = it == mat_by_name.end() ? -1 : it->second;
    } else if (tag == "mtllib") {
      std::string name;
      ss >> name;
      obj_load_mtl(folder + name, out.materials, mat_by_name);
    }
  }
  for (auto& t : tris)
    for (auto& c : t) {
      bool hn = c.n >= 0 && (size_t)c.n < N.size() / 3, ht = c.t >= 0 && (size_t)c.t < T.size() / 2;
      any_n |= hn; any_t |= ht; all_n &= hn; all_t &= ht;
    }
  bool use_n = any_n && all_n, use_t = any_t && all_t;
  uint32_t index = 0;
  for (size_t f = 0; f < tris.size(); f++) {
    for (auto& c : tris[f]) {
      out.v.insert(out.v.end(), {P[3 * c.v], P[3 * c.v + 1], P[3 * c.v + 2]});
      if (use_n) out.n.insert(out.n.end(), {N[3 * c.n], N[3 * c.n + 1], N[3 * c.n + 2]});
      if (use_t) out.uv.insert(out.uv.end(), {T[2 * c.t], T[2 * c.t + 1]});
    }
    out.idx.insert(out.idx.end(), {index, index + 1, index + 2});
    index += 3;
    out.face_material.push_back(tri_mat[f]);
  }
  if (tris.empty()) {
    err = "no faces";
    return false;
  }
  return true;
}
