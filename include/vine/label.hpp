// vine/label.hpp -- the user-replaceable functor API of the solver.
//
// Same names and meaning as the reference's compile-time label / constraint /
// edge types for CSP == 1 (include/label.hpp:170-247, include/constraint.hpp:81-105,
// include/edge.hpp:16-26), as __host__ __device__ functors so the same type
// serves the HIP kernels and host code.  The solver kernels are templates over
// <Label_t, Constraint_t, Edge_t>; a user-defined label drops in by providing
// the same members plus a label_traits specialisation (how the label's two
// 32-bit criteria are packed into the 8-byte Pareto-store entry).
//
// Differences from the reference, each deliberate:
//  - expand_cpu is additive like expand (the reference's host twin multiplies,
//    include/label.hpp:224-232, which keeps every host label at (0,0));
//  - functors take const references; labels carry no vtable.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#else
// plain host compilers: the functors are ordinary inline members
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
struct float2 {
  float x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
#endif

namespace vine {

// c_edge (include/edge.hpp:16-26): the two arc weights.
struct c_edge {
  float len;
  float delay;
  __host__ __device__ c_edge() : len(0.f), delay(0.f) {}
  __host__ __device__ c_edge(float l, float d) : len(l), delay(d) {}
};

// c_label (include/label.hpp:170-247): vertex id plus the two path sums.
struct c_label {
  int vid;
  float d;
  float delay;

  __host__ __device__ c_label() : vid(-1), d(0.f), delay(0.f) {}
  // source label: both sums zero (include/label.hpp:184-189)
  __host__ __device__ explicit c_label(int v) : vid(v), d(0.f), delay(0.f) {}
  __host__ __device__ c_label(int v, float d_, float dl) : vid(v), d(d_), delay(dl) {}

  __host__ __device__ static bool equal(const c_label* a, const c_label* b) {
    return a->vid == b->vid && a->d == b->d && a->delay == b->delay;
  }
  __host__ __device__ static void copy(c_label* a, const c_label* b) { *a = *b; }
  __host__ __device__ int get_vid() const { return vid; }

  // son = this extended by arc e into next_v (include/label.hpp:213-222)
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, c_label* son) const {
    son->vid = next_v;
    son->d = d + e.len;
    son->delay = delay + e.delay;
  }
  template <typename Edge_t>
  void expand_cpu(int next_v, const Edge_t& e, c_label* son) const {
    expand(next_v, e, son);
  }

  // non-strict dominance (include/label.hpp:234-236)
  __host__ __device__ static bool dominate(const c_label& a, const c_label& b) {
    return a.d <= b.d && a.delay <= b.delay;
  }
  __host__ __device__ static bool dominate_vid(const c_label& a, const c_label& b) {
    return a.vid == b.vid && dominate(a, b);
  }
  static bool dominate_cpu(const c_label& a, const c_label& b) { return dominate(a, b); }
};

// c_constraint (include/constraint.hpp:81-105): delay bound, inclusive.
template <typename Label_t>
struct c_constraint {
  float delay_L;
  __host__ __device__ c_constraint() : delay_L(0.f) {}
  __host__ __device__ explicit c_constraint(float d) : delay_L(d) {}
  void set_constraints(int d) { delay_L = (float)d; }
  __host__ __device__ bool pass_constraints_check(const Label_t& l) const {
    return !(l.delay > delay_L);
  }
  bool pass_constraints_check_cpu(const Label_t& l) const { return pass_constraints_check(l); }
};

// How a label's criteria live in the 8-byte Pareto-store entry, and the
// scheduling key (the criterion the frontier window orders by).
template <typename Label_t>
struct label_traits;

template <>
struct label_traits<c_label> {
  __host__ __device__ static float2 pack(const c_label& l) { return make_float2(l.d, l.delay); }
  __host__ __device__ static c_label unpack(int vid, float2 p) { return c_label(vid, p.x, p.y); }
  __host__ __device__ static float key(const c_label& l) { return l.d; }
};

}  // namespace vine
