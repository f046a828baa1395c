// vine/families.hpp -- the reference's other label families, as functors.
//
// The reference selects one label/constraint pair at compile time with CSP
// (include/label.hpp:46-169, 248-461; include/constraint.hpp:21-80, 106-186).
// Here every family is available at once, each in its own namespace with the
// reference's names (c_label, c_constraint<Label_t>, expand, dominate, copy,
// get_vid, pass_constraints_check), and the generic label-correcting kernels
// (vine_amd/csrc/families.hip) are templates over them.  c_edge is the shared
// arc type of vine/label.hpp.
//
//   family (vine.h)           reference     label fields        expand               dominate (a over b)                  constraint
//   VINE_FAMILY_HOP           CSP == 2      d, delay, hop       +len, +delay, +1     d<=, delay<=, hop<=                  delay <= L, hop <= H
//   VINE_FAMILY_HOP_MAXCOST   CSP == 3      + maxcost           maxcost = max(.,len) + maxcost<=                          + maxcost <= M
//   VINE_FAMILY_2OSP          CSP == 4      d, delay            +len, +delay         d<=, delay<=                         delay <= L
//   VINE_FAMILY_SOCIAL        CSP == 5      d, delay (from 1)   *len, *delay         d>=, delay>=                         delay <= L
//   VINE_FAMILY_QOS           default       d, bw (from 21474836) +len, min(bw,delay) d<=, bw>=                          bw >= L
//
// Not in the reference (it never extracts an answer for these families):
// family_traits<Label_t> names the answer -- the best label of t's Pareto set
// in a fixed lexicographic order -- the four output fields, the family id of
// the C ABI, and how a constraint maps to and from the per-query limits
// (delay_L, hop_L, maxcost_L).
// Defaults hop_L = 100 and maxcost_L = 10000 are the reference's constructor
// defaults (include/constraint.hpp:31, 61).
#pragma once

#include "vine.h"
#include "vine/label.hpp"

namespace vine {

// ---- CSP == 2: hop-bounded (include/label.hpp:46-105, constraint.hpp:21-49)
namespace hop {
struct c_label {
  int vid;
  float d;
  float delay;
  int hop;
  __host__ __device__ c_label() : vid(-1), d(0.f), delay(0.f), hop(0) {}
  __host__ __device__ explicit c_label(int v) : vid(v), d(0.f), delay(0.f), hop(0) {}
  __host__ __device__ static void copy(c_label* a, const c_label* b) { *a = *b; }
  __host__ __device__ int get_vid() const { return vid; }
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, c_label* son) const {
    son->vid = next_v;
    son->d = d + e.len;
    son->delay = delay + e.delay;
    son->hop = hop + 1;
  }
  __host__ __device__ static bool dominate(const c_label& a, const c_label& b) {
    return a.d <= b.d && a.delay <= b.delay && a.hop <= b.hop;
  }
};
template <typename Label_t>
struct c_constraint {
  float delay_L;
  int hop_L;
  __host__ __device__ c_constraint() : delay_L(0.f), hop_L(100) {}
  __host__ __device__ c_constraint(float d, int h = 100) : delay_L(d), hop_L(h) {}
  void set_constraints(float d, int h) {
    delay_L = d;
    hop_L = h;
  }
  __host__ __device__ bool pass_constraints_check(const Label_t& l) const {
    return !(l.delay > delay_L || l.hop > hop_L);
  }
};
}  // namespace hop

// ---- CSP == 3: hop-bounded with a bottleneck arc length (include/label.hpp:106-169,
// constraint.hpp:50-80)
namespace hop_maxcost {
struct c_label {
  int vid;
  float d;
  float delay;
  int hop;
  float maxcost;
  __host__ __device__ c_label() : vid(-1), d(0.f), delay(0.f), hop(0), maxcost(0.f) {}
  __host__ __device__ explicit c_label(int v) : vid(v), d(0.f), delay(0.f), hop(0), maxcost(0.f) {}
  __host__ __device__ static void copy(c_label* a, const c_label* b) { *a = *b; }
  __host__ __device__ int get_vid() const { return vid; }
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, c_label* son) const {
    son->vid = next_v;
    son->d = d + e.len;
    son->delay = delay + e.delay;
    son->hop = hop + 1;
    son->maxcost = maxcost > e.len ? maxcost : e.len;
  }
  __host__ __device__ static bool dominate(const c_label& a, const c_label& b) {
    return a.d <= b.d && a.delay <= b.delay && a.hop <= b.hop && a.maxcost <= b.maxcost;
  }
};
template <typename Label_t>
struct c_constraint {
  float delay_L;
  int hop_L;
  float maxcost_L;
  __host__ __device__ c_constraint() : delay_L(0.f), hop_L(100), maxcost_L(10000.f) {}
  __host__ __device__ c_constraint(float d, int h = 100, float m = 10000.f) : delay_L(d), hop_L(h), maxcost_L(m) {}
  void set_constraints(float d, int h, int m) {
    delay_L = d;
    hop_L = h;
    maxcost_L = (float)m;
  }
  __host__ __device__ bool pass_constraints_check(const Label_t& l) const {
    return !(l.delay > delay_L || l.hop > hop_L || l.maxcost > maxcost_L);
  }
};
}  // namespace hop_maxcost

// ---- CSP == 4: 2-OSP, the same label and constraint as CSP == 1
// (include/label.hpp:248-303, constraint.hpp:106-133)
namespace osp2 {
using c_label = vine::c_label;
template <typename Label_t>
using c_constraint = vine::c_constraint<Label_t>;
}  // namespace osp2

// ---- CSP == 5: social, multiplicative, larger is better (include/label.hpp:304-381,
// constraint.hpp:134-161).  Sources start at (1, 1).
namespace social {
struct c_label {
  int vid;
  float d;
  float delay;
  __host__ __device__ c_label() : vid(-1), d(1.f), delay(1.f) {}
  __host__ __device__ explicit c_label(int v) : vid(v), d(1.f), delay(1.f) {}
  __host__ __device__ static bool equal(const c_label* a, const c_label* b) {
    return a->vid == b->vid && a->d == b->d && a->delay == b->delay;
  }
  __host__ __device__ static void copy(c_label* a, const c_label* b) { *a = *b; }
  __host__ __device__ int get_vid() const { return vid; }
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, c_label* son) const {
    son->vid = next_v;
    son->d = d * e.len;
    son->delay = delay * e.delay;
  }
  __host__ __device__ static bool dominate(const c_label& a, const c_label& b) {
    return a.d >= b.d && a.delay >= b.delay;
  }
  __host__ __device__ static bool dominate_vid(const c_label& a, const c_label& b) {
    return a.vid == b.vid && dominate(a, b);
  }
};
template <typename Label_t>
struct c_constraint {
  float delay_L;
  __host__ __device__ c_constraint() : delay_L(0.f) {}
  __host__ __device__ explicit c_constraint(float d) : delay_L(d) {}
  void set_constraints(int d) { delay_L = (float)d; }
  __host__ __device__ bool pass_constraints_check(const Label_t& l) const { return !(l.delay > delay_L); }
};
}  // namespace social

// ---- default branch: network QoS, additive length and bottleneck bandwidth
// (include/label.hpp:382-461, constraint.hpp:162-186).  `delay` carries the
// bandwidth: the minimum arc delay along the path, sources start at 21474836.
namespace qos {
struct c_label {
  int vid;
  float d;
  float delay;
  __host__ __device__ c_label() : vid(-1), d(0.f), delay(21474836.f) {}
  __host__ __device__ explicit c_label(int v) : vid(v), d(0.f), delay(21474836.f) {}
  __host__ __device__ static bool equal(const c_label* a, const c_label* b) {
    return a->vid == b->vid && a->d == b->d && a->delay == b->delay;
  }
  __host__ __device__ static void copy(c_label* a, const c_label* b) { *a = *b; }
  __host__ __device__ int get_vid() const { return vid; }
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, c_label* son) const {
    son->vid = next_v;
    son->d = d + e.len;
    son->delay = delay < e.delay ? delay : e.delay;
  }
  __host__ __device__ static bool dominate(const c_label& a, const c_label& b) {
    return a.d <= b.d && a.delay >= b.delay;
  }
  __host__ __device__ static bool dominate_vid(const c_label& a, const c_label& b) {
    return a.vid == b.vid && dominate(a, b);
  }
};
template <typename Label_t>
struct c_constraint {
  float delay_L;
  __host__ __device__ c_constraint() : delay_L(0.f) {}
  __host__ __device__ explicit c_constraint(float d) : delay_L(d) {}
  void set_constraints(int d) { delay_L = (float)d; }
  __host__ __device__ bool pass_constraints_check(const Label_t& l) const { return !(l.delay < delay_L); }
};
}  // namespace qos

// ---- answer order, output fields and constraints from limits (not in the reference)
template <typename Label_t>
struct family_traits;

// lexicographic "a before b" on up to four keys, each ascending (+1) or descending (-1)
__host__ __device__ inline bool lex_before(const float* a, const float* b, const int* dir, int k) {
  for (int i = 0; i < k; i++) {
    if (a[i] != b[i]) return dir[i] > 0 ? a[i] < b[i] : a[i] > b[i];
  }
  return false;
}

template <>
struct family_traits<vine::c_label> {  // CSP == 1 and 2-OSP: min (d, delay)
  using constraint = vine::c_constraint<vine::c_label>;
  static constexpr int family = VINE_FAMILY_CSP;
  static void limits(const constraint& C, float* lim) {
    lim[0] = C.delay_L;
    lim[1] = 0.f;
    lim[2] = 0.f;
  }
  __host__ __device__ static void fields(const vine::c_label& l, float* f) {
    f[0] = l.d;
    f[1] = l.delay;
    f[2] = 0.f;
    f[3] = 0.f;
  }
  __host__ __device__ static bool better(const vine::c_label& a, const vine::c_label& b) {
    return a.d < b.d || (a.d == b.d && a.delay < b.delay);
  }
  __host__ __device__ static constraint make(const float* lim) { return constraint(lim[0]); }
};

template <>
struct family_traits<hop::c_label> {  // min (d, delay, hop)
  using constraint = hop::c_constraint<hop::c_label>;
  static constexpr int family = VINE_FAMILY_HOP;
  static void limits(const constraint& C, float* lim) {
    lim[0] = C.delay_L;
    lim[1] = (float)C.hop_L;
    lim[2] = 0.f;
  }
  __host__ __device__ static void fields(const hop::c_label& l, float* f) {
    f[0] = l.d;
    f[1] = l.delay;
    f[2] = (float)l.hop;
    f[3] = 0.f;
  }
  __host__ __device__ static bool better(const hop::c_label& a, const hop::c_label& b) {
    const float x[3] = {a.d, a.delay, (float)a.hop}, y[3] = {b.d, b.delay, (float)b.hop};
    const int dir[3] = {1, 1, 1};
    return lex_before(x, y, dir, 3);
  }
  __host__ __device__ static constraint make(const float* lim) { return constraint(lim[0], (int)lim[1]); }
};

template <>
struct family_traits<hop_maxcost::c_label> {  // min (d, delay, hop, maxcost)
  using constraint = hop_maxcost::c_constraint<hop_maxcost::c_label>;
  static constexpr int family = VINE_FAMILY_HOP_MAXCOST;
  static void limits(const constraint& C, float* lim) {
    lim[0] = C.delay_L;
    lim[1] = (float)C.hop_L;
    lim[2] = C.maxcost_L;
  }
  __host__ __device__ static void fields(const hop_maxcost::c_label& l, float* f) {
    f[0] = l.d;
    f[1] = l.delay;
    f[2] = (float)l.hop;
    f[3] = l.maxcost;
  }
  __host__ __device__ static bool better(const hop_maxcost::c_label& a, const hop_maxcost::c_label& b) {
    const float x[4] = {a.d, a.delay, (float)a.hop, a.maxcost}, y[4] = {b.d, b.delay, (float)b.hop, b.maxcost};
    const int dir[4] = {1, 1, 1, 1};
    return lex_before(x, y, dir, 4);
  }
  __host__ __device__ static constraint make(const float* lim) {
    return constraint(lim[0], (int)lim[1], lim[2]);
  }
};

template <>
struct family_traits<social::c_label> {  // max (d, delay)
  using constraint = social::c_constraint<social::c_label>;
  static constexpr int family = VINE_FAMILY_SOCIAL;
  static void limits(const constraint& C, float* lim) {
    lim[0] = C.delay_L;
    lim[1] = 0.f;
    lim[2] = 0.f;
  }
  __host__ __device__ static void fields(const social::c_label& l, float* f) {
    f[0] = l.d;
    f[1] = l.delay;
    f[2] = 0.f;
    f[3] = 0.f;
  }
  __host__ __device__ static bool better(const social::c_label& a, const social::c_label& b) {
    return a.d > b.d || (a.d == b.d && a.delay > b.delay);
  }
  __host__ __device__ static constraint make(const float* lim) { return constraint(lim[0]); }
};

template <>
struct family_traits<qos::c_label> {  // min d, then max bandwidth
  using constraint = qos::c_constraint<qos::c_label>;
  static constexpr int family = VINE_FAMILY_QOS;
  static void limits(const constraint& C, float* lim) {
    lim[0] = C.delay_L;
    lim[1] = 0.f;
    lim[2] = 0.f;
  }
  __host__ __device__ static void fields(const qos::c_label& l, float* f) {
    f[0] = l.d;
    f[1] = l.delay;
    f[2] = 0.f;
    f[3] = 0.f;
  }
  __host__ __device__ static bool better(const qos::c_label& a, const qos::c_label& b) {
    return a.d < b.d || (a.d == b.d && a.delay > b.delay);
  }
  __host__ __device__ static constraint make(const float* lim) { return constraint(lim[0]); }
};

}  // namespace vine
