// pt_trace.hpp -- two-level BVH4 closest-hit traversal (replaces OptiX's
// rtTrace: programs/raygen.cu:77, programs/materials/light_sample.cuh:69).
//
// One ray per lane.  The traversal stack keeps its top entries in LDS (one
// column per lane, bank-conflict free; how many is the kernel's choice) and
// spills deeper entries to a per-thread slice of HBM.  Children are visited
// nearest first by entry distance (ties by slot); popped entries farther
// than the current hit are skipped.  Candidates win on smaller t, ties on the smaller
// (instance, primitive) key, so the result does not depend on visit order.
#pragma once

#include "pt_device.hpp"

namespace pt {

template <int NL>
struct StackT {
  uint32_t* ref;  // LDS column, stride BLOCK
  float* tn;
  uint2* spill;   // this thread's HBM slice (STK_TOTAL entries)
  int sp;
  uint32_t* overflow;
  PTD void put(int pos, uint32_t r, float t) {
    if (pos < NL) {
      ref[pos * BLOCK] = r;
      tn[pos * BLOCK] = t;
    } else if (pos < STK_TOTAL) {
      spill[pos - NL] = make_uint2(r, __float_as_uint(t));
    } else {
      atomicAdd(overflow, 1u);
    }
  }
  PTD void pop(uint32_t& r, float& t) {
    sp--;
    if (sp < NL) {
      r = ref[sp * BLOCK];
      t = tn[sp * BLOCK];
    } else if (sp < STK_TOTAL) {
      uint2 e = spill[sp - NL];
      r = e.x;
      t = __uint_as_float(e.y);
    } else {
      r = EMPTY_CHILD;
      t = FINF;
    }
  }
};
using Stack = StackT<STK_LDS>;

PTD float4 as_f4(uint4 u) {
  return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
}

struct HitOut {
  float t;
  int inst, prim;  // original instance id and primitive index (-1: miss)
  int pos;         // instance position in top-level order
};

struct TraceCounts {
  uint32_t nodes, prims, insts;
};

// entry distance of one child slot or +inf
PTD float slab_test(float lox, float hix, float loy, float hiy, float loz, float hiz, v3 o, v3 inv, float tmin,
                    float tmax) {
  float t0x = (lox - o.x) * inv.x, t1x = (hix - o.x) * inv.x;
  float t0y = (loy - o.y) * inv.y, t1y = (hiy - o.y) * inv.y;
  float t0z = (loz - o.z) * inv.z, t1z = (hiz - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
  float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
  return (tn <= tf * SLAB_FAR_SCALE) ? tn : FINF;
}

// EXT: the scene holds instances of the extended kinds (volumes,
// cylinders, moving spheres); volume free flights are drawn from `vseed`
// (the ray's RNG state) and the volume and instance ids, moving spheres are
// placed at `time`.
// KINDS: bit k set when instances of PT_GEOM kind k (spheres, rects,
// boxes, meshes) may occur; the tests of the other kinds (and the
// triangle test's per-instance ray setup) are compiled out.
constexpr int KIND_ALL = 15;
// The traversal of one ray as a resumable state: begin() sets it up, each
// step() does one trip of the loop (a pop, one internal node, one leaf's
// primitives or an instance entry) and returns false once the ray is done
// (h then holds the result).  trace_closest runs the steps back to back;
// kernels that refill lanes with new rays step several rays at once.
//
// ANY && any_hit: the ray ends at the first candidate accepted in
// (tmin, tmax) instead of the closest (occlusion queries whose answer is
// only whether anything is hit: a miss is a miss either way, and a hit is
// reported as one).  ANY is a template switch so that the closest-hit
// kernels carry no trace of the early exit (a runtime-false flag still
// changes their traversal loop's code and costs them about a tenth of
// their time).
template <bool COUNT, bool EXT = false, int KINDS = KIND_ALL, bool ANY = false>
struct Tracer {
  static constexpr bool SPH = (KINDS & 1) != 0, RECT = (KINDS & 2) != 0, BOX = (KINDS & 4) != 0,
                        MESH = (KINDS & 8) != 0;
  HitOut h;
  v3 o, d, winv, ro, rd, rinv;
  float tmin, time;
  uint32_t vseed;
  TriRay tr;
  int cur;      // instance slot in top-level order, -1 = top level
  int base_sp;  // stack depth when the instance was entered
  uint32_t node_base, prim_base, kind, inst_id;
  // the entry in hand: the top-level root first; the nearest child of each
  // node and an instance's root are taken directly instead of through the
  // stack, which visits the same nodes in the same order
  uint32_t ref;
  float etn;
  bool have, any_hit;
  // bound on loop trips: a well-formed structure never comes close
  uint32_t trips;

  // the watertight test of one triangle record and, when it is nearer (or
  // as near with the smaller key), its acceptance; true when an any-hit
  // query ends here
  PTD bool offer_tri(float4 a, float4 b, float4 c, TraceCounts& cnt) {
    if (COUNT) cnt.prims++;
    float t = FINF, b1, b2;
    if (!tri_test(tr, ro, mk3(a), mk3(b), mk3(c), t, b1, b2)) t = FINF;
    const uint32_t orig = __float_as_uint(a.w);
    if (!(t > tmin)) return false;
    bool take = t < h.t;
    if (!take && t == h.t && h.inst >= 0)
      take = ((int)inst_id < h.inst) || ((int)inst_id == h.inst && (int)orig < h.prim);
    if (!take) return false;
    h.t = t;
    h.inst = (int)inst_id;
    h.prim = (int)orig;
    h.pos = cur;
    return ANY && any_hit;
  }

  template <class ST>
  PTD void begin(const DevScene& sc, v3 o_, v3 d_, float tmin_, float tmax, ST& st, uint32_t vseed_ = 0u,
                 float time_ = 0.f, bool any_hit_ = false) {
    h = HitOut{tmax, -1, -1, -1};
    o = o_;
    d = d_;
    tmin = tmin_;
    time = time_;
    vseed = vseed_;
    any_hit = any_hit_;
    winv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    ro = o;
    rd = d;
    rinv = winv;
    tr = TriRay{0, 1, 2, 0.f, 0.f, 0.f};
    cur = -1;
    base_sp = 0;
    node_base = sc.tlas_root;
    prim_base = 0;
    kind = 0;
    inst_id = 0;
    st.sp = 0;
    ref = 0u;
    etn = tmin;
    have = true;
    trips = 0;
  }

  template <class ST>
  PTD bool step(const DevScene& sc, ST& st, TraceCounts& cnt) {
    if (!(++trips < (1u << 22))) return false;
    if (!have) {
      if (cur >= 0 && st.sp == base_sp) {
        cur = -1;
        node_base = sc.tlas_root;
        ro = o; rd = d; rinv = winv;
      }
      if (st.sp == 0) return false;
      st.pop(ref, etn);
      if (ref == EMPTY_CHILD || etn > h.t) return true;
    }
    have = false;
    if (ref & LEAF_BIT) {
      uint32_t first = ref & 0x0FFFFFFFu;
      uint32_t count = ((ref >> 28) & 7u) + 1u;
      if (cur < 0) {
        // enter the instance at top-level position `first`
        const uint4* ir = sc.insts + (size_t)first * 8;
        float4 w0 = as_f4(ir[0]);
        float4 w1 = as_f4(ir[1]);
        float4 w2 = as_f4(ir[2]);
        uint4 info = ir[3];
        if (COUNT) cnt.insts++;
        node_base = info.x;
        prim_base = info.y;
        kind = info.z & 0xFFu;
        inst_id = info.w;
        // the object-space ray is formed for every instance and kept where
        // the instance has a transform: formed behind a branch on that flag,
        // the matrix would be loaded only after the flag arrives (a second
        // memory round trip per instance entry)
        // (component selects: a select of whole structs keeps this state
        // in scratch memory)
        const bool xf = (info.z >> 8) != 0;
        const v3 xo = xf_point(w0, w1, w2, o), xd = xf_vector(w0, w1, w2, d);
        ro = mk3(xf ? xo.x : o.x, xf ? xo.y : o.y, xf ? xo.z : o.z);
        rd = mk3(xf ? xd.x : d.x, xf ? xd.y : d.y, xf ? xd.z : d.z);
        if (xf) rinv = mk3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        if (MESH && kind == PT_GEOM_MESH) tr = tri_setup(rd);
        cur = (int)first;
        base_sp = st.sp;
        ref = 0u;  // the bottom-level root, next
        have = true;
        return true;
      }
      if (MESH && kind == PT_GEOM_MESH) {
        // triangles two at a time: both records are loaded before either
        // test, one memory round trip per pair (a leaf holds up to 4); the
        // candidates are still offered in leaf order
        for (uint32_t k = 0; k < count; k += 2) {
          const uint32_t p0 = prim_base + first + k;
          const bool two = k + 1 < count;
          const uint32_t p1 = two ? p0 + 1 : p0;
          const float4 a0 = sc.tri[3 * p0], b0 = sc.tri[3 * p0 + 1], c0 = sc.tri[3 * p0 + 2];
          const float4 a1 = sc.tri[3 * p1], b1 = sc.tri[3 * p1 + 1], c1 = sc.tri[3 * p1 + 2];
          if (offer_tri(a0, b0, c0, cnt)) return false;
          if (two && offer_tri(a1, b1, c1, cnt)) return false;
        }
        return true;
      }
      for (uint32_t k = 0; k < count; k++) {
        uint32_t p = prim_base + first + k;
        if (COUNT) cnt.prims++;
        float t = FINF;
        uint32_t orig = 0xFFFFFFFFu;
        if (SPH && kind == PT_GEOM_SPHERES) {
          t = isect_sphere(sc.sph[p], ro, rd, tmin);
        } else if (RECT && kind == PT_GEOM_RECTS) {
          float4 A = sc.rect[2 * p], B = sc.rect[2 * p + 1];
          t = isect_rect(A, B.x, (int)__float_as_uint(B.y), ro, rd, rinv);
          orig = __float_as_uint(B.w);
        } else if (BOX && kind == PT_GEOM_BOXES) {
          float4 lo = sc.box[2 * p], hi = sc.box[2 * p + 1];
          t = isect_box(mk3(lo), mk3(hi), ro, rinv, tmin);
          orig = __float_as_uint(lo.w);
        } else if (EXT && kind == PT_GEOM_VOLUMES) {
          float4 A = sc.vol[2 * p], B = sc.vol[2 * p + 1];
          orig = __float_as_uint(B.w) & 0x7FFFFFFFu;
          t = isect_volume(A, B, ro, rd, tmin, vol_xi(vseed, orig, inst_id));
        } else if (EXT && kind == PT_GEOM_CYLINDERS) {
          float4 A = sc.cyl[2 * p], B = sc.cyl[2 * p + 1];
          orig = __float_as_uint(B.w);
          t = isect_cylinder(A, B, ro, rd, tmin);
        } else if (EXT && kind == PT_GEOM_MOVING_SPHERES) {
          float4 A = sc.mov[3 * p], B = sc.mov[3 * p + 1], C = sc.mov[3 * p + 2];
          v3 c = moving_center(A, B, C, time);
          orig = __float_as_uint(B.w);
          t = isect_sphere(make_float4(c.x, c.y, c.z, A.w), ro, rd, tmin);
        } else if (MESH) {
          float4 a = sc.tri[3 * p], b = sc.tri[3 * p + 1], c = sc.tri[3 * p + 2];
          float b1, b2;
          if (!tri_test(tr, ro, mk3(a), mk3(b), mk3(c), t, b1, b2)) t = FINF;
          orig = __float_as_uint(a.w);
        }
        if (!(t > tmin)) continue;
        bool take = t < h.t;
        if (!take && t == h.t && h.inst >= 0) {
          if (SPH && kind == PT_GEOM_SPHERES) orig = sc.sph_idx[p];
          take = ((int)inst_id < h.inst) || ((int)inst_id == h.inst && (int)orig < h.prim);
        }
        if (take) {
          if (SPH && kind == PT_GEOM_SPHERES && orig == 0xFFFFFFFFu) orig = sc.sph_idx[p];
          h.t = t;
          h.inst = (int)inst_id;
          h.prim = (int)orig;
          h.pos = cur;
          if (ANY && any_hit) return false;
        }
      }
      return true;
    }
    // internal node
    if (COUNT) cnt.nodes++;
    const uint4* nd = sc.nodes + (size_t)(node_base + ref) * 8;
    uint4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3], q4 = nd[4], q5 = nd[5], q6 = nd[6], q7 = nd[7];
    const uint32_t lox[4] = {q0.x, q0.y, q0.z, q0.w}, hix[4] = {q1.x, q1.y, q1.z, q1.w};
    const uint32_t loy[4] = {q2.x, q2.y, q2.z, q2.w}, hiy[4] = {q3.x, q3.y, q3.z, q3.w};
    const uint32_t loz[4] = {q4.x, q4.y, q4.z, q4.w}, hiz[4] = {q5.x, q5.y, q5.z, q5.w};
    const uint32_t chd[4] = {q6.x, q6.y, q6.z, q6.w}, cnc[4] = {q7.x, q7.y, q7.z, q7.w};
    float tn[4];
    uint32_t cref[4];
    // every slot's test runs and empty slots are masked afterwards: a test
    // behind a branch on the slot lets the compiler load that slot's planes
    // only after the child words arrive, a second memory round trip per node
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const float ts = slab_test(__uint_as_float(lox[c]), __uint_as_float(hix[c]), __uint_as_float(loy[c]),
                                 __uint_as_float(hiy[c]), __uint_as_float(loz[c]), __uint_as_float(hiz[c]), ro, rinv,
                                 tmin, h.t);
      tn[c] = chd[c] != EMPTY_CHILD ? ts : FINF;
      cref[c] = cnc[c] ? (LEAF_BIT | ((cnc[c] - 1u) << 28) | chd[c]) : chd[c];
    }
    // rank by (tn, slot): the nearest is taken next, the others go on the
    // stack with the second nearest on top
    int nh = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) nh += (tn[c] != FINF);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (tn[c] == FINF) continue;
      int rank = 0;
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (e != c && tn[e] != FINF && (tn[e] < tn[c] || (tn[e] == tn[c] && e < c))) rank++;
      if (rank == 0) {
        // the cull a pop would apply (the slab test's far bound is 2^-22 loose)
        if (!(tn[c] > h.t)) {
          ref = cref[c];
          etn = tn[c];
          have = true;
        }
      } else {
        st.put(st.sp + nh - 1 - rank, cref[c], tn[c]);
      }
    }
    if (nh > 0) st.sp += nh - 1;
    return true;
  }
};

template <bool COUNT, bool EXT = false, int KINDS = KIND_ALL, bool ANY = false, class ST>
PTD HitOut trace_closest(const DevScene& sc, v3 o, v3 d, float tmin, float tmax, ST& st, TraceCounts& cnt,
                         uint32_t vseed = 0u, float time = 0.f, bool any_hit = false) {
  Tracer<COUNT, EXT, KINDS, ANY> t;
  t.begin(sc, o, d, tmin, tmax, st, vseed, time, any_hit);
  while (t.step(sc, st, cnt)) {
  }
  return t.h;
}

}  // namespace pt
