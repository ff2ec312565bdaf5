// zf_membrane_dev.h -- membrane force terms written once over the scalar
// type T: T = double is the force evaluation, T = Dual<N> evaluates the
// same expressions in forward-mode automatic differentiation (value plus N
// directional derivatives), which the membrane adjoint uses for exact local
// Jacobian-vector products.  The order of operations is the reference's
// (ZebraflowWT.c:5800-5831, :6042-6086, :6888-7384, :7527-7531).
#ifndef ZF_MEMBRANE_DEV_H
#define ZF_MEMBRANE_DEV_H

#include <type_traits>
#include <cmath>

#include "zf_internal.h"

namespace zf {
namespace memb {

struct MembParams {
  double L, kT, kbend, kd, ka, kv, area0, vol0, x_max, x_floor, cos_cut, area_lo, area_hi, cos_blend;
  double force_limit, vol_limit, sliver_len, sliver_len_min, sliver_f1, sliver_f2, relax, relax2;
};

inline MembParams params_of(const zf_membrane_config &m) {
  return MembParams{m.len_scale, m.kT, m.kbend, m.kd, m.ka, m.kv, m.area0, m.vol0, m.x_max,
                    m.x_floor, m.cos_cut, m.area_lo, m.area_hi, m.cos_blend, m.force_limit, m.vol_limit,
                    m.sliver_len, m.sliver_len_min, m.sliver_f1, m.sliver_f2, m.relax, m.relax2};
}

// ---- forward-mode dual numbers -------------------------------------------------
template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ Dual() {}
  __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
};

__device__ __forceinline__ double val(double x) { return x; }
template <int N>
__device__ __forceinline__ double val(const Dual<N> &x) { return x.v; }

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N> &b) { return b + a; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] / b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a / b.v;
  const double k = -r.v / b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = k * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double k = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = k * a.d[i];
  return r;
}
using ::sqrt;  // the double overloads, hidden by the one above
template <int N>
__device__ __forceinline__ bool operator<(const Dual<N> &a, double b) { return a.v < b; }
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N> &a, double b) { return a.v > b; }
template <int N>
__device__ __forceinline__ bool operator>=(const Dual<N> &a, double b) { return a.v >= b; }
template <int N>
__device__ __forceinline__ bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }

// ---- 3-vectors ------------------------------------------------------------------
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
__device__ __forceinline__ V3<T> operator-(const V3<T> &a, const V3<T> &b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
__device__ __forceinline__ V3<T> operator+(const V3<T> &a, const V3<T> &b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
__device__ __forceinline__ V3<T> operator*(double s, const V3<T> &a) {
  return {s * a.x, s * a.y, s * a.z};
}
template <int N>
__device__ __forceinline__ V3<Dual<N>> operator*(const Dual<N> &s, const V3<Dual<N>> &a) {
  return {s * a.x, s * a.y, s * a.z};
}
template <class T>
__device__ __forceinline__ T dot(const V3<T> &a, const V3<T> &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
__device__ __forceinline__ V3<T> cross(const V3<T> &a, const V3<T> &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// The type of the constants in arithmetic of type T: FP32 with FP32, FP64
// with FP64 and the dual numbers (so those keep their expressions exactly).
template <class T>
struct ScalarOf {
  using type = double;
};
template <>
struct ScalarOf<float> {
  using type = float;
};
__device__ __forceinline__ V3<float> operator*(float s, const V3<float> &a) {
  return {s * a.x, s * a.y, s * a.z};
}
template <class T>
__device__ __forceinline__ V3<T> lift(const V3<double> &a) {
  return {T(a.x), T(a.y), T(a.z)};
}
__device__ __forceinline__ V3<double> ld3(const double *a, size_t stride, size_t i) {
  return {a[i], a[stride + i], a[2 * stride + i]};
}
// A position in the arithmetic type: FP32 works relative to the node the
// force is for (the difference taken in FP64, so no absolute coordinate is
// rounded to FP32); FP64 and the dual numbers take it as it is.
template <class T>
__device__ __forceinline__ V3<T> liftp(const V3<double> &a, const V3<double> &o) {
  if constexpr (std::is_same<T, float>::value) return lift<T>(a - o);
  else return lift<T>(a);
}
template <class T>
__device__ __forceinline__ int nbad(const V3<T> &f, double lim) {
  auto b = [lim](double v) { return (isnan(v) || isinf(v) || fabs(v) > lim) ? 1 : 0; };
  return b(val(f.x)) + b(val(f.y)) + b(val(f.z));
}

// ---- triangle geometry (ZebraflowWT.c:5800-5831) --------------------------------
// A = (e1 x e3)/2 with e1 = n2-n1, e3 = n1-n3; outward normal -A/|A|.
template <class T>
struct TriG {
  T a;
  V3<T> m, C;
};
template <class T>
__device__ __forceinline__ TriG<T> tri_geom(const V3<T> &p1, const V3<T> &p2, const V3<T> &p3, double L) {
  TriG<T> g;
  g.C = (1.0 / 3.0) * (p1 + p2 + p3);
  const V3<T> e1 = p2 - p1, e3 = p1 - p3;
  const V3<T> A = (0.5 * L * L) * cross(e1, e3);
  g.a = sqrt(dot(A, A));
  g.m = {-A.x / g.a, -A.y / g.a, -A.z / g.a};
  return g;
}
// Stored geometry (area, normal, centroid) of triangle q, as constants;
// the centroid as liftp takes positions (relative to o in FP32).
template <class T>
__device__ __forceinline__ TriG<T> tri_load(const double *__restrict__ tg, size_t TT, size_t q,
                                            const V3<double> &o = V3<double>{0, 0, 0}) {
  TriG<T> g;
  g.a = T(tg[q]);
  g.m = {T(tg[TT + q]), T(tg[2 * TT + q]), T(tg[3 * TT + q])};
  g.C = liftp<T>(V3<double>{tg[4 * TT + q], tg[5 * TT + q], tg[6 * TT + q]}, o);
  return g;
}

// ---- one rhombus (edge) for one of its four nodes ----------------------------------
struct EdgeConst {
  double lmax, plen, kp, s0, c0, ae1, ae2;
  bool own1, own2;
};

template <class T>
struct EdgeOut {
  V3<T> f;   // spring + bending + local and global area terms (after the guards)
  V3<T> fb;  // the bending part of f
  V3<T> fv;  // volume gradient (1/6)((a/3) m + lc x d) of the owned triangles
  int fails;
  bool has_fv;
};

// Role r = 0..3 is node n1..n4 of the rhombus (n1, n2: the edge; n3, n4:
// the wing tips of triangles t1, t2), ZebraflowWT.c:6888-7384.  Within the
// edge the terms are summed in the reference's order: spring, bending,
// tri1 local+global area, tri2.
template <class T>
__device__ __forceinline__ EdgeOut<T> edge_role(int r, const V3<T> &p1, const V3<T> &p2,
                                                const V3<T> &p3, const V3<T> &p4, const TriG<T> &g1,
                                                const TriG<T> &g2, const EdgeConst &ec,
                                                const V3<T> &cc, const T &f_ag,
                                                const MembParams &mp) {
  using S = typename ScalarOf<T>::type;  // the constants' type: FP32 in FP32 arithmetic
  const S L = S(mp.L);
  EdgeOut<T> o;
  o.f = {T(0.0), T(0.0), T(0.0)};
  o.fb = o.f;
  o.fv = o.f;
  o.fails = 0;
  o.has_fv = false;
  // WLC + POW spring (:6888-6924)
  const V3<T> d12 = p1 - p2;
  T len = sqrt(dot(d12, d12));
  const V3<T> e1v = (S(1.0) / len) * d12;
  T ratio = len * L / S(ec.lmax);
  if (ratio > S(mp.x_max)) ratio = T(mp.x_max);
  if (ratio < S(mp.x_floor)) len = T(mp.x_floor * ec.lmax / mp.L);
  const T wlc = S(-mp.kT / ec.plen) * (ratio - S(0.25) + S(1.0) / (S(4.0) * (1 - ratio) * (1 - ratio)));
  const T comp = S(ec.kp) / (len * len * L * L);
  if (r == 0) o.f = (S(mp.relax2) * wlc) * e1v + (S(mp.relax2) * comp) * e1v;
  if (r == 1) o.f = (S(-mp.relax2) * wlc) * e1v + (S(-mp.relax2) * comp) * e1v;
  // clamped triangle areas (:6929-6951)
  T a1 = g1.a, a2 = g2.a;
  if (a1 < mp.area_lo * ec.ae1) a1 = T(mp.area_lo * ec.ae1);
  else if (a1 > mp.area_hi * ec.ae1) a1 = T(mp.area_hi * ec.ae1);
  if (a2 < mp.area_lo * ec.ae2) a2 = T(mp.area_lo * ec.ae2);
  else if (a2 > mp.area_hi * ec.ae2) a2 = T(mp.area_hi * ec.ae2);
  const V3<T> &m1 = g1.m, &m2 = g2.m;
  T ct = dot(m1, m2);
  if (ct > 1.0) ct = T(1.0);
  else if (ct < -1.0) ct = T(-1.0);
  const T chk = dot(m1 - m2, L * (g1.C - g2.C));
  const T st = chk >= 0.0 ? sqrt(1 - ct * ct) : -sqrt(1 - ct * ct);
  // bending (:6984-7081)
  T bb = fabs(val(ct)) > mp.cos_cut ? T(0.0) : S(mp.kbend) * (st * S(ec.c0) - ct * S(ec.s0)) / st;
  if (mp.cos_blend > 0.0 && fabs(val(ct)) <= mp.cos_cut && fabs(val(ct)) > mp.cos_cut - mp.cos_blend) {
    // smooth cut-off (no counterpart in the reference): bb times the
    // smoothstep t^2 (3 - 2t) of t = (cos_cut - |cos|)/cos_blend, C1 in the
    // positions, equal to the reference's bb below cos_cut - cos_blend
    const T ac = val(ct) < 0.0 ? -ct : ct;
    const T t = (T(mp.cos_cut) - ac) * S(1.0 / mp.cos_blend);
    bb = bb * (t * t * (T(3.0) - S(2.0) * t));
  }
  const V3<T> dp21 = L * (p2 - p1), dp24 = L * (p2 - p4), dp32 = L * (p3 - p2), dp41 = L * (p4 - p1),
              dp13 = L * (p1 - p3);
  const V3<T> cp1 = a1 * cross(m1, dp21), cp2 = a2 * cross(m2, dp21), cp3 = a1 * cross(m1, dp24),
              cp4 = a2 * cross(m2, dp24), cp5 = a1 * cross(m1, dp32), cp6 = a2 * cross(m2, dp32),
              cp7 = a1 * cross(m1, dp41), cp8 = a2 * cross(m2, dp41), cp9 = a1 * cross(m1, dp13),
              cp10 = a2 * cross(m2, dp13);
  const T b11 = -bb * ct / (a1 * a1), b12 = bb / (a1 * a2), b22 = -bb * ct / (a2 * a2);
  V3<T> fb;
  if (r == 0) fb = b11 * cp5 + b12 * (cp3 + cp6) + b22 * cp4;
  else if (r == 1) fb = b11 * cp9 + b12 * (cp7 + cp10) + b22 * cp8;
  else if (r == 2) fb = b11 * cp1 + b12 * cp2;
  else fb = (-b12) * cp1 + (-b22) * cp2;
  const int kb = nbad(fb, mp.force_limit);
  o.fails += kb;
  if (!kb) {
    o.f = o.f + fb;
    o.fb = fb;
  }
  // area and volume terms of the triangles this edge owns (:7126-7359)
  for (int side = 0; side < 2; side++) {
    if (side == 0 && (!ec.own1 || r == 3)) continue;
    if (side == 1 && (!ec.own2 || r == 2)) continue;
    const T &a = side ? a2 : a1;
    const double ae = side ? ec.ae2 : ec.ae1;
    const V3<T> &m = side ? m2 : m1;
    const V3<T> lc = L * ((side ? g2.C : g1.C) - cc);
    V3<T> g, gv;
    if (side == 0) {
      g = r == 0 ? cp5 : r == 1 ? cp9 : cp1;
      gv = r == 0 ? cross(lc, dp32) : r == 1 ? cross(lc, dp13) : cross(lc, dp21);
    } else {
      g = r == 0 ? cp4 : r == 1 ? cp8 : S(-1.0) * cp2;
      gv = r == 0 ? cross(lc, dp24) : r == 1 ? cross(lc, dp41) : S(-1.0) * cross(lc, dp21);
    }
    const T alpha = S(-mp.kd * mp.relax) * (a - S(ae)) / (S(4.0 * ae) * a);
    const V3<T> fa = alpha * g;
    const V3<T> faa = (S(0.25) * f_ag / a) * g;
    const V3<T> fv = S(1.0 / 6.0) * ((a / S(3.0)) * m + gv);
    const int ka = nbad(fa, mp.force_limit);
    o.fails += ka;
    if (!ka) o.f = o.f + fa;
    const int kaa = nbad(faa, mp.force_limit);
    o.fails += kaa;
    if (!kaa) o.f = o.f + faa;
    if (!nbad(fv, mp.vol_limit)) {
      o.fv = o.fv + fv;
      o.has_fv = true;
    }
  }
  return o;
}

// Anti-sliver force on node xn from one of its triangles (centroid ct),
// ZebraflowWT.c:6042-6086.
template <class T>
__device__ __forceinline__ V3<T> sliver(const V3<T> &xn, const V3<T> &ct, const MembParams &mp) {
  const V3<T> ev = xn - ct;
  const T l = sqrt(dot(ev, ev));
  if (!(l < mp.sliver_len)) return {T(0.0), T(0.0), T(0.0)};
  using S = typename ScalarOf<T>::type;
  const T mag = l > mp.sliver_len_min ? S(mp.relax2 * mp.sliver_f1) / l : T(mp.relax2 * mp.sliver_f2);
  return (mag / l) * ev;
}

// Mesh tables of the shared RBC mesh (device pointers).
struct MeshTables {
  const int32_t *tri;       // 3 x nt, 0-based
  const int32_t *node_tri;  // nn x 10: count, triangles (0-based)
  const int32_t *inc_start, *inc;  // node -> (edge << 2 | role), edge order
  const int32_t *E;         // 8 x ne: n1 n2 n3 n4 t1 t2 own1 own2 (0-based)
  const double *EQ;         // 5 x ne: lmax plength kp sin0 cos0
  const double *tae;        // rest areas, nt
  int nn, nt, ne;
};

__device__ __forceinline__ EdgeConst edge_const(const MeshTables &mt, int e) {
  const int ne = mt.ne;
  EdgeConst ec;
  ec.lmax = mt.EQ[e];
  ec.plen = mt.EQ[ne + e];
  ec.kp = mt.EQ[2 * ne + e];
  ec.s0 = mt.EQ[3 * ne + e];
  ec.c0 = mt.EQ[4 * ne + e];
  ec.ae1 = mt.tae[mt.E[4 * ne + e]];
  ec.ae2 = mt.tae[mt.E[5 * ne + e]];
  ec.own1 = mt.E[6 * ne + e] != 0;
  ec.own2 = mt.E[7 * ne + e] != 0;
  return ec;
}

template <class T>
struct NodeForce {
  V3<T> f;   // F_Total of the node
  V3<T> fb;  // its bending part
  V3<double> nrm;
  int fails;
};

// F_Total of node n of the cell whose nodes start at nbase and triangles at
// tbase, from the stored triangle geometry; the cell scalars (centroid cc,
// global area force f_ag, volume force f_vs) are of type T, cc relative to
// the node in FP32 (liftp).  SPLIT selects the role with a switch (for
// wide T, fewer live registers).
template <class T, bool SPLIT = false>
__device__ __forceinline__ NodeForce<T> node_force(int n, size_t nbase, size_t tbase,
                                                   const double *__restrict__ X, size_t NT,
                                                   const double *__restrict__ tg, size_t TT,
                                                   const MeshTables &mt, const V3<T> &cc,
                                                   const T &f_ag, const T &f_vs,
                                                   const MembParams &mp) {
  NodeForce<T> out;
  const V3<double> org = ld3(X, NT, nbase + n);
  const V3<T> xn = liftp<T>(org, org);
  // node area and normal, anti-sliver terms, in triangle order
  V3<T> f = {T(0.0), T(0.0), T(0.0)}, fbs = f;
  V3<double> nrm = {0, 0, 0};
  double narea = 0;
  const int ntri = mt.node_tri[10 * n];
  for (int k = 1; k <= ntri && k < 10; k++) {
    const int t = mt.node_tri[10 * n + k];
    const size_t qt = tbase + t;
    const double a = tg[qt];
    const V3<double> mv = {tg[TT + qt], tg[2 * TT + qt], tg[3 * TT + qt]};
    narea += a / 3.0;
    nrm = nrm + (a / 3.0) * mv;
    const V3<T> ct = liftp<T>(V3<double>{tg[4 * TT + qt], tg[5 * TT + qt], tg[6 * TT + qt]}, org);
    f = f + sliver(xn, ct, mp);
  }
  const double inv = 1.0 / narea;
  out.nrm = {nrm.x * inv, nrm.y * inv, nrm.z * inv};
  // every rhombus containing the node, in edge order
  V3<T> fv = {T(0.0), T(0.0), T(0.0)};
  int fails = 0;
  const int ne = mt.ne;
  for (int k = mt.inc_start[n]; k < mt.inc_start[n + 1]; k++) {
    const int code = mt.inc[k], e = code >> 2;
    const V3<T> p1 = liftp<T>(ld3(X, NT, nbase + mt.E[e]), org),
                p2 = liftp<T>(ld3(X, NT, nbase + mt.E[ne + e]), org),
                p3 = liftp<T>(ld3(X, NT, nbase + mt.E[2 * ne + e]), org),
                p4 = liftp<T>(ld3(X, NT, nbase + mt.E[3 * ne + e]), org);
    const TriG<T> g1 = tri_load<T>(tg, TT, tbase + mt.E[4 * ne + e], org);
    const TriG<T> g2 = tri_load<T>(tg, TT, tbase + mt.E[5 * ne + e], org);
    const EdgeConst ec = edge_const(mt, e);
    EdgeOut<T> o;
    if (SPLIT) {  // one inlined copy per role: only that role's terms stay live
      switch (code & 3) {
        case 0: o = edge_role(0, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        case 1: o = edge_role(1, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        case 2: o = edge_role(2, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        default: o = edge_role(3, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp);
      }
    } else {
      o = edge_role(code & 3, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp);
    }
    f = f + o.f;
    fbs = fbs + o.fb;
    if (o.has_fv) fv = fv + o.fv;
    fails += o.fails;
  }
  // volume constraint (:7527-7531)
  f = f + (typename ScalarOf<T>::type(mp.relax) * f_vs) * fv;
  out.f = f;
  out.fb = fbs;
  out.fails = fails;
  return out;
}

// Host side (zf_membrane.hip): the mesh tables of a context, and the
// triangle geometry (7 x ncap*nt) and cell scalars (8 x ncap, as the
// context's) of positions X, on the context stream.
MeshTables mesh_tables(const zf_ctx *ctx);
int membrane_geometry(zf_ctx *ctx, const double *X, double *tg, double *cell);

}  // namespace memb
}  // namespace zf

#endif
