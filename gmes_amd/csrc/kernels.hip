// CDNA4 kernels of the FDTD engine.
//
// Data layout in HBM: each Yee component is its own C-order array on the
// reference's staggered storage shape (gmes/geometry.py:276-358), complex
// values interleaved (re, im) like numpy.  A component's cells carry a
// uint16 material id; the id selects a record (class + coefficients) from
// a small table.  CPML/UPML cells keep their auxiliary words in full-grid
// arrays of the component; dispersive (ADE) cells live in compact
// per-class lists with structure-of-arrays state.
//
// Kernels per leapfrog step (gmes/fdtd.py:1099-1122):
//   k_ghost<H>  periodic/Bloch ghost planes of H     (gmes/fdtd.py:955-1097)
//   k_main<E>   curl + Dielectric/CPML/UPML/Const    (src/pw_{dielectric,cpml,upml,const}.hh)
//   k_ade_*     Drude/Lorentz, DCP-ADE, DCP-PLRC, DM2 (src/pw_{drude,lorentz,dcp,dm2}.hh)
//   k_src<E>    point sources, TFSF face corrections (gmes/pw_source.py)
//   k_probe<E>
//   (auxiliary 1-D grids)
//   k_ghost<E>, k_main<H>, k_src<H>, k_probe<H>
// Every FP64 expression keeps the reference's operation order (division by
// d1/d2, dt/eps first, left-to-right sums) and the build disables FMA
// contraction, so the FP64 path is bit-identical to the reference's.
#include "engine_kernels.h"

namespace gmes_engine {

// ---------------------------------------------------------------- arithmetic
template <class R> __device__ __forceinline__ R vadd(R a, R b) { return a + b; }
template <class R> __device__ __forceinline__ R vsub(R a, R b) { return a - b; }
template <class R> __device__ __forceinline__ R vscl(R s, R a) { return s * a; }
template <class R> __device__ __forceinline__ R vdiv(R a, R d) { return a / d; }
template <class R> __device__ __forceinline__ cplx<R> vadd(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R> __device__ __forceinline__ cplx<R> vsub(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R> __device__ __forceinline__ cplx<R> vscl(R s, cplx<R> a) { return {s * a.re, s * a.im}; }
template <class R> __device__ __forceinline__ cplx<R> vdiv(cplx<R> a, R d) { return {a.re / d, a.im / d}; }
template <class R> __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> __device__ __forceinline__ T vzero() { return T{}; }

// Global-memory accessors.  A device allocation reached through a pointer
// read from memory (a batch job table) is not known to be global: plain
// accesses through it become generic (flat) instructions, which also wait
// on LDS traffic.  These access it through an address-space-1 pointer.
#define GMES_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const GMES_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(GMES_GLOBAL T*)p = v;
}
// complex words and records move as native vectors (a class-typed copy
// would take a generic pointer again)
template <class R>
using vec2_t = R __attribute__((ext_vector_type(2)));
typedef int i4_t __attribute__((ext_vector_type(4)));
template <class R>
__device__ __forceinline__ cplx<R> gld(const cplx<R>* p) {
  const vec2_t<R> v = *(const GMES_GLOBAL vec2_t<R>*)p;
  return {v.x, v.y};
}
template <class R>
__device__ __forceinline__ void gst(cplx<R>* p, cplx<R> v) {
  *(GMES_GLOBAL vec2_t<R>*)p = vec2_t<R>{v.re, v.im};
}
template <class R>
__device__ __forceinline__ DevRec<R> load_rec(const DevRec<R>* p) {
  constexpr int n = sizeof(DevRec<R>) / 16;
  union {
    DevRec<R> r;
    i4_t v[n];
  } u;
#pragma unroll
  for (int c = 0; c < n; ++c) u.v[c] = ((const GMES_GLOBAL i4_t*)p)[c];
  return u.r;
}

// Streaming accesses: field words and material ids are touched once per
// launch, so they are loaded and stored non-temporally (no reuse to keep
// in L2).  GMES_NT=0 builds the plain form.
#ifndef GMES_NT
#define GMES_NT 1
#endif
template <class R> __device__ __forceinline__ R ld_nt(const R* p) {
#if GMES_NT
  return __builtin_nontemporal_load((const __attribute__((address_space(1))) R*)p);
#else
  return *(const __attribute__((address_space(1))) R*)p;
#endif
}
template <class R> __device__ __forceinline__ cplx<R> ld_nt(const cplx<R>* p) {
  return {ld_nt(&p->re), ld_nt(&p->im)};
}
template <class R> __device__ __forceinline__ void st_nt(R* p, R v) {
#if GMES_NT
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) R*)p);
#else
  *(__attribute__((address_space(1))) R*)p = v;
#endif
}
template <class R> __device__ __forceinline__ void st_nt(cplx<R>* p, cplx<R> v) {
  st_nt(&p->re, v.re);
  st_nt(&p->im, v.im);
}
template <class R> __device__ __forceinline__ R re_of(R a) { return a; }
template <class R> __device__ __forceinline__ R re_of(cplx<R> a) { return a.re; }
template <class R> __device__ __forceinline__ R im_of(R) { return R(0); }
template <class R> __device__ __forceinline__ R im_of(cplx<R> a) { return a.im; }
template <class T> __device__ __forceinline__ T make_val(typename real_of<T>::type re, typename real_of<T>::type im);
template <> __device__ __forceinline__ float make_val<float>(float re, float) { return re; }
template <> __device__ __forceinline__ double make_val<double>(double re, double) { return re; }
template <> __device__ __forceinline__ cplx<float> make_val<cplx<float>>(float re, float im) { return {re, im}; }
template <> __device__ __forceinline__ cplx<double> make_val<cplx<double>>(double re, double im) { return {re, im}; }


template <class T>
__device__ __forceinline__ T ld(const FieldArgs<T>& a, int c, int i, int j, int k) {
  // an inactive component is (1,1,1) and reads as zero (src/cpp23_support.hh:52-74)
  if (!(a.active & (1u << c))) return vzero<T>();
  return gld(a.f[c] + didx(a, c, i, j, k));
}

// Curl differences D1, D2 of component C at (i,j,k): src/pw_dielectric.hh
// :118-119 (Ex) :157-158 (Ey) :197-198 (Ez) :292-293 (Hx) :331-332 (Hy) :370-371 (Hz)
template <int C, class T>
__device__ __forceinline__ void curl_terms(const FieldArgs<T>& a, int i, int j, int k, T& D1, T& D2) {
  if constexpr (C == 0) {
    D1 = vsub(ld(a, 5, i + 1, j + 1, k), ld(a, 5, i + 1, j, k));
    D2 = vsub(ld(a, 4, i + 1, j, k + 1), ld(a, 4, i + 1, j, k));
  } else if constexpr (C == 1) {
    D1 = vsub(ld(a, 3, i, j + 1, k + 1), ld(a, 3, i, j + 1, k));
    D2 = vsub(ld(a, 5, i + 1, j + 1, k), ld(a, 5, i, j + 1, k));
  } else if constexpr (C == 2) {
    D1 = vsub(ld(a, 4, i + 1, j, k + 1), ld(a, 4, i, j, k + 1));
    D2 = vsub(ld(a, 3, i, j + 1, k + 1), ld(a, 3, i, j, k + 1));
  } else if constexpr (C == 3) {
    D1 = vsub(ld(a, 2, i, j, k - 1), ld(a, 2, i, j - 1, k - 1));
    D2 = vsub(ld(a, 1, i, j - 1, k), ld(a, 1, i, j - 1, k - 1));
  } else if constexpr (C == 4) {
    D1 = vsub(ld(a, 0, i - 1, j, k), ld(a, 0, i - 1, j, k - 1));
    D2 = vsub(ld(a, 2, i, j, k - 1), ld(a, 2, i - 1, j, k - 1));
  } else {
    D1 = vsub(ld(a, 1, i, j - 1, k), ld(a, 1, i - 1, j - 1, k));
    D2 = vsub(ld(a, 0, i - 1, j, k), ld(a, 0, i - 1, j - 1, k));
  }
}

template <int C> struct Ax;
template <> struct Ax<0> { static constexpr int a1 = 1, a2 = 2, a0 = 0; };
template <> struct Ax<1> { static constexpr int a1 = 2, a2 = 0, a0 = 1; };
template <> struct Ax<2> { static constexpr int a1 = 0, a2 = 1, a0 = 2; };
template <> struct Ax<3> { static constexpr int a1 = 1, a2 = 2, a0 = 0; };
template <> struct Ax<4> { static constexpr int a1 = 2, a2 = 0, a0 = 1; };
template <> struct Ax<5> { static constexpr int a1 = 0, a2 = 1, a0 = 2; };

// ------------------------------------------------------------- main update
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ cplx<float> fma_(float a, cplx<float> b, cplx<float> c) {
  return {__builtin_fmaf(a, b.re, c.re), __builtin_fmaf(a, b.im, c.im)};
}

// PRE: the CPML psi words pre1, pre2 (or the UPML state word pre1, for the
// components in m.upml_mask) of a real-FP32 cell were loaded with the
// cell's other operands (prefetch_psi); otherwise update_value loads them.
//
// CPML psi words that can be nonzero.  Along an axis where the profile's
// c coefficient is exactly zero (outside that axis's layers: sigma = 0),
// psi starts at +0 and stays +0, since b*(+0) + (+-0)*x = +0; computing it
// from a literal +0 gives the same value, and the word is neither read nor
// written.  For a CPML record the engine packs, per axis, the coordinate
// run [lo, hi) where c is zero for every component as lo | hi << 16, into
// n1 (x), n2 (y) and coef_off (z), fields CPML records do not otherwise use.
template <int A, class R>
__device__ __forceinline__ bool psi_live(const DevRec<R>& rec, const int ix[3]) {
  const int run = A == 0 ? rec.n1 : A == 1 ? rec.n2 : (int)rec.coef_off;
  return ix[A] < (run & 0xFFFF) || ix[A] >= (run >> 16);
}

// New value of a Dielectric, CPML or UPML cell of component C from its old
// value f and curl differences D1, D2.  st is the component's CPML/UPML
// state array (same layout as the field), flat the cell's element.
//   FP64 keeps the reference's operation order (division by d1/d2, dt/eps
//   first, left-to-right sums): src/pw_dielectric.hh:118,292,
//   src/pw_cpml.hh:127-132,339-344, src/pw_upml.hh:124-128,328-332.
//   FP32 uses reciprocals the host computed in FP64 (1/d, 1/kappa in the
//   CPML profile, 1/eps in the record) and explicit FMAs.
// CPML profile rows (b, c, 1/kappa) of the cell's coordinate along x, y, z,
// loaded with the other operands (MainArgs::pp); c = 0, 1/kappa = 1 inside
// the zero-c run of the axis.
struct PRows {
  float v[3][3];
  // UPML (MainArgs::ur): per axis, c1..c4 of the row the two components
  // with that curl axis share, and c5, c6 of the row of the component
  // that owns the axis
  float us[3][4];
  float uo[3][2];
};

template <int C, class T, bool PRE, bool PP = false, bool UR = false>
__device__ __forceinline__ T update_value(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>& rec,
                                          T* __restrict__ st, int flat, int i, int j, int k, T f, T D1, T D2,
                                          T pre1, T pre2, const PRows& pr) {
  using R = typename real_of<T>::type;
  constexpr bool E = C < 3;
  const int cls = rec.cls;
  if constexpr (std::is_same<R, float>::value) {
    const float id1 = m.inv_d[Ax<C>::a1], id2 = m.inv_d[Ax<C>::a2];
    const T a = vscl(id1, D1), b = vscl(id2, D2);  // D1/d1, D2/d2
    if (cls == GMES_CLS_DIELECTRIC) return E ? fma_(rec.c0, vsub(a, b), f) : fma_(rec.c0, vsub(b, a), f);
    const int ix[3] = {i, j, k};
    // profile rows of set 0 come from the arguments (no dependent load)
    const GMES_GLOBAL int64_t* po = (const GMES_GLOBAL int64_t*)m.prof_off + (rec.prof * 6 + C) * 3;
    const bool set0 = rec.prof == 0;
    if (cls == GMES_CLS_CPML) {
      float b1, c1, k1, b2, c2, k2;
      if (PP && set0) {
        b1 = pr.v[Ax<C>::a1][0]; c1 = pr.v[Ax<C>::a1][1]; k1 = pr.v[Ax<C>::a1][2];
        b2 = pr.v[Ax<C>::a2][0]; c2 = pr.v[Ax<C>::a2][1]; k2 = pr.v[Ax<C>::a2][2];
      } else {
        const GMES_GLOBAL float* p1 = (const GMES_GLOBAL float*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a1] : po[Ax<C>::a1]) + ix[Ax<C>::a1] * 3;
        const GMES_GLOBAL float* p2 = (const GMES_GLOBAL float*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a2] : po[Ax<C>::a2]) + ix[Ax<C>::a2] * 3;
        b1 = p1[0]; c1 = p1[1]; k1 = p1[2];
        b2 = p2[0]; c2 = p2[1]; k2 = p2[2];
      }
      const bool u1 = psi_live<Ax<C>::a1>(rec, ix), u2 = psi_live<Ax<C>::a2>(rec, ix);
      T psi1, psi2;
      if constexpr (sizeof(T) >= 8) {
        const T o1 = u1 ? gld(st + flat) : vzero<T>();
        const T o2 = u2 ? gld(st + m.dsize[C] + flat) : vzero<T>();
        psi1 = fma_(c1, a, vscl(b1, o1));
        psi2 = fma_(c2, b, vscl(b2, o2));
        if (u1) st_nt(st + flat, psi1);
        if (u2) st_nt(st + m.dsize[C] + flat, psi2);
      } else {
        // real FP32: a dead psi (coordinate inside the zero-c run of its
        // axis) is neither loaded nor stored; it is +0 in memory and in the
        // formula.  With the profile rows prefetched, masking these lanes
        // beat reading a guard word and storing +0 back: 256^3 E launch
        // 0.181 to 0.171 ms.
        const T o1 = PRE ? pre1 : (u1 ? gld(st + flat) : T{});
        const T o2 = PRE ? pre2 : (u2 ? gld(st + m.dsize[C] + flat) : T{});
        psi1 = fma_(c1, a, vscl(b1, o1));
        psi2 = fma_(c2, b, vscl(b2, o2));
        if (u1) st_nt(st + flat, psi1);  // used once per launch, like the field words
        if (u2) st_nt(st + m.dsize[C] + flat, psi2);
      }
      const T t = vsub(vadd(fma_(-k2, b, vscl(k1, a)), psi1), psi2);
      return E ? fma_(rec.c0, t, f) : fma_(-rec.c0, t, f);
    }
    // UPML: c1, c2 along a1, c3, c4 along a2, c5, c6 along the own axis
    float u1, u2, u3, u4, u5, u6;
    if (UR && set0) {
      u1 = pr.us[Ax<C>::a1][0]; u2 = pr.us[Ax<C>::a1][1];
      u3 = pr.us[Ax<C>::a2][2]; u4 = pr.us[Ax<C>::a2][3];
      u5 = pr.uo[Ax<C>::a0][0]; u6 = pr.uo[Ax<C>::a0][1];
    } else {
      const GMES_GLOBAL float* p1 = (const GMES_GLOBAL float*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a1] : po[Ax<C>::a1]) + ix[Ax<C>::a1] * 6;
      const GMES_GLOBAL float* p2 = (const GMES_GLOBAL float*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a2] : po[Ax<C>::a2]) + ix[Ax<C>::a2] * 6;
      const GMES_GLOBAL float* p0 = (const GMES_GLOBAL float*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a0] : po[Ax<C>::a0]) + ix[Ax<C>::a0] * 6;
      u1 = p1[0]; u2 = p1[1]; u3 = p2[2]; u4 = p2[3]; u5 = p0[4]; u6 = p0[5];
    }
    const T dstore = (PRE && ((m.upml_mask >> C) & 1u)) ? pre1 : gld(st + flat);
    const T cu = vsub(a, b);
    const T dn = E ? fma_(u2, cu, vscl(u1, dstore)) : fma_(-u2, cu, vscl(u1, dstore));
    st_nt(st + flat, dn);
    return fma_(u4 * rec.inv_eps, fma_(-u6, dstore, vscl(u5, dn)), vscl(u3, f));
  } else {
    const R d1 = m.d[Ax<C>::a1], d2 = m.d[Ax<C>::a2];
    if (cls == GMES_CLS_DIELECTRIC) {
      if constexpr (E) return vadd(f, vscl(rec.c0, vsub(vdiv(D1, d1), vdiv(D2, d2))));
      else return vadd(f, vscl(rec.c0, vsub(vdiv(D2, d2), vdiv(D1, d1))));
    }
    const int ix[3] = {i, j, k};
    // profile rows of set 0 come from the arguments (no dependent load)
    const GMES_GLOBAL int64_t* po = (const GMES_GLOBAL int64_t*)m.prof_off + (rec.prof * 6 + C) * 3;
    const bool set0 = rec.prof == 0;
    if (cls == GMES_CLS_CPML) {
      const GMES_GLOBAL R* p1 = (const GMES_GLOBAL R*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a1] : po[Ax<C>::a1]) + ix[Ax<C>::a1] * 3;
      const GMES_GLOBAL R* p2 = (const GMES_GLOBAL R*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a2] : po[Ax<C>::a2]) + ix[Ax<C>::a2] * 3;
      const bool u1 = psi_live<Ax<C>::a1>(rec, ix), u2 = psi_live<Ax<C>::a2>(rec, ix);
      const T o1 = u1 ? gld(st + flat) : vzero<T>();
      const T o2 = u2 ? gld(st + m.dsize[C] + flat) : vzero<T>();
      const T psi1 = vadd(vscl(p1[0], o1), vdiv(vscl(p1[1], D1), d1));
      const T psi2 = vadd(vscl(p2[0], o2), vdiv(vscl(p2[1], D2), d2));
      if (u1) st_nt(st + flat, psi1);
      if (u2) st_nt(st + m.dsize[C] + flat, psi2);
      const T t = vsub(vadd(vsub(vdiv(vdiv(D1, d1), p1[2]), vdiv(vdiv(D2, d2), p2[2])), psi1), psi2);
      if constexpr (E) return vadd(f, vscl(rec.c0, t));
      else return vsub(f, vscl(rec.c0, t));
    }
    // UPML
    const GMES_GLOBAL R* p1 = (const GMES_GLOBAL R*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a1] : po[Ax<C>::a1]) + ix[Ax<C>::a1] * 6;
    const GMES_GLOBAL R* p2 = (const GMES_GLOBAL R*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a2] : po[Ax<C>::a2]) + ix[Ax<C>::a2] * 6;
    const GMES_GLOBAL R* p0 = (const GMES_GLOBAL R*)m.prof + (set0 ? (int64_t)m.po0[C][Ax<C>::a0] : po[Ax<C>::a0]) + ix[Ax<C>::a0] * 6;
    const T dstore = gld(st + flat);
    const T cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
    T dn;
    if constexpr (E) dn = vadd(vscl(p1[0], dstore), vscl(p1[1], cu));
    else dn = vsub(vscl(p1[0], dstore), vscl(p1[1], cu));
    st_nt(st + flat, dn);
    return vadd(vscl(p2[2], f), vdiv(vscl(p2[3], vsub(vscl(p0[4], dn), vscl(p0[5], dstore))), rec.eps_mu));
  }
}

// Update of one cell of component C: Dummy (no-op), Const (src/pw_const.hh
// :96,198), Dielectric, CPML, UPML.  ADE classes are left to their kernels.
// f is the cell's old value, loaded by the caller.
template <int C, class T, bool PRE = false, bool PP = false, bool UR = false>
__device__ __forceinline__ T store_cell(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>& rec,
                                        T* __restrict__ F, T* __restrict__ st, int flat, int i, int j, int k,
                                        T f, T D1, T D2, T pre1 = T{}, T pre2 = T{}, const PRows& pr = PRows{}) {
  const int cls = rec.cls;
  if (cls == GMES_CLS_DUMMY || cls >= GMES_CLS_DRUDE) return f;
  if (cls == GMES_CLS_CONST) {
    const T v = make_val<T>(rec.val_re, rec.val_im);
    st_nt(F + flat, v);
    return v;
  }
  const T v = update_value<C, T, PRE, PP, UR>(m, rec, st, flat, i, j, k, f, D1, D2, pre1, pre2, pr);
  st_nt(F + flat, v);
  return v;
}

// Periodic / Bloch ghost planes of E written by the thread that owns the
// source cell (the copy k_ghost<E> makes, fused into the E update when
// nothing else writes or reads the planes involved before that copy would
// run).  x is the cell's new value; ix its storage index.  Same phase
// products and corner order as k_ghost: ph1 * x on the a1 ghost plane,
// ph2 * x on the a2 plane, ph2 * (ph1 * x) on the corner line.
template <class T>
__device__ __forceinline__ T ghost_phase(const cplx<typename real_of<T>::type>& ph, T x) {
  if constexpr (is_cplx<T>::value) return cmul(ph, x);
  else return x;
}

template <class T>
__device__ __forceinline__ void fused_ghost(const MainArgs<T>& m, int C, T* __restrict__ F, const int ix[3], T x) {
  const int a1 = m.gax[C][0], a2 = m.gax[C][1];
  const int* sh = m.fa.s[C];
  const int dst1 = sh[a1] - 1, dst2 = sh[a2] - 1;
  const bool on1 = ix[a1] == 0, on2 = ix[a2] == 0;
  if (!on1 && !on2) return;
  int g[3] = {ix[0], ix[1], ix[2]};
  T x1 = x;
  if (on1) {
    x1 = ghost_phase<T>(m.gph[C][0], x);
    g[a1] = dst1;
    gst(F + didx(m.fa, C, g[0], g[1], g[2]), x1);
    g[a1] = 0;
  }
  if (on2) {
    g[a2] = dst2;
    gst(F + didx(m.fa, C, g[0], g[1], g[2]), ghost_phase<T>(m.gph[C][1], x));
    if (on1) {
      g[a1] = dst1;
      gst(F + didx(m.fa, C, g[0], g[1], g[2]), ghost_phase<T>(m.gph[C][1], x1));
    }
  }
}

template <int C, class T>
__device__ __forceinline__ void update_cell(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>& rec,
                                            int flat, int i, int j, int k, T D1, T D2) {
  const int cls = rec.cls;
  if (cls == GMES_CLS_DUMMY || cls >= GMES_CLS_DRUDE) return;
  T* F = m.fa.f[C];
  (void)store_cell<C>(m, rec, F, static_cast<T*>(m.pml[C]), flat, i, j, k,
                      cls == GMES_CLS_CONST ? vzero<T>() : gld(F + flat), D1, D2);
}

template <int C, class T>
__device__ __forceinline__ void main_cell(const MainArgs<T>& m, int i, int j, int k) {
  const FieldArgs<T>& a = m.fa;
  const int flat = didx(a, C, i, j, k);
  const DevRec<typename real_of<T>::type> rec = load_rec(m.recs + gld(m.ids[C] + flat));
  if (rec.cls == GMES_CLS_DUMMY || rec.cls >= GMES_CLS_DRUDE) return;
  T D1 = vzero<T>(), D2 = vzero<T>();
  if (rec.cls != GMES_CLS_CONST) curl_terms<C>(a, i, j, k, D1, D2);
  update_cell<C>(m, rec, flat, i, j, k, D1, D2);
}

// Generic main update (any mode): one thread per (i,j,k); every component
// of the field type whose storage contains (i,j,k) is updated.
template <bool E, class T>
__device__ __forceinline__ void main_generic_cell(const MainArgs<T>& m, int i, int j, int k) {
  const FieldArgs<T>& a = m.fa;
  constexpr int C0 = E ? 0 : 3;
#define GMES_MAIN_COMP(CC)                                                              \
  if ((a.active & (1u << (CC))) && i < a.s[CC][0] && j < a.s[CC][1] && k < a.s[CC][2]) \
    main_cell<CC>(m, i, j, k);
  GMES_MAIN_COMP(C0 + 0)
  GMES_MAIN_COMP(C0 + 1)
  GMES_MAIN_COMP(C0 + 2)
#undef GMES_MAIN_COMP
}

// kloop: rows of at most 2 elements (a collapsed z axis, as in TMz/TEz):
// threads run along j and each walks the few k of its row, so no lane sits
// on a row position that does not exist or holds only ghosts.
template <bool E, class T>
__global__ void __launch_bounds__(256) k_main(MainArgs<T> m) {
  const int i = blockIdx.z, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (m.kloop) {
    for (int k = 0; k < m.kext; ++k) main_generic_cell<E, T>(m, i, j, k);
  } else {
    main_generic_cell<E, T>(m, i, j, blockIdx.x * blockDim.x + threadIdx.x);
  }
}

// Record table staging: the table is copied as 16-byte chunks, one per
// thread, loaded into registers first and written to LDS later.
template <class R>
struct RecStage {
  static constexpr int kChunks = sizeof(DevRec<R>) / 16;
  static_assert(sizeof(DevRec<R>) % 16 == 0, "record size must be a multiple of 16 bytes");
  i4_t v;
  int t;
  bool use_lds;
  __device__ __forceinline__ void load(const DevRec<R>* g, int nrec, int tid, int nthreads) {
    t = tid;
    use_lds = nrec * kChunks <= nthreads && nrec <= kRecLds;
    if (use_lds && t < nrec * kChunks) v = gld(reinterpret_cast<const i4_t*>(g) + t);
  }
  __device__ __forceinline__ void commit(DevRec<R>* lds, int nrec) {
    if (!use_lds) return;
    if (t < nrec * kChunks) reinterpret_cast<i4_t*>(lds)[t] = v;
    __syncthreads();
  }
};

// 2-D modes with a collapsed z axis (TMz, TEz): every row holds one or two
// elements, so the generic kernel's threads (one per element) would each
// run one short dependent chain (id, record, curl operands, update).  Here
// a thread owns (i, j) for N consecutive i and every component of the
// field type at its one updated z plane (k = 0 for E and for rows of one
// element, k = 1 for the H components whose k = 0 plane is the ghost).  All
// ids, old values and curl operands of its cells are loaded first; ghost
// cells, whose neighbours lie outside the arrays, load no operands.  The
// record table is staged into LDS meanwhile, and only then do the updates
// run.
template <class T, int N>
struct Cells2z {
  T f[3][N], d1[3][N], d2[3][N];
  int flat[3][N];
  uint16_t id[3][N];
  bool upd[3][N];  // an in-range, non-ghost cell of an active component
};

// Ghost cells are the last plane (E) or the first plane (H) along each
// transverse axis whose extent exceeds 1.
template <int C, class T>
__device__ __forceinline__ bool ghost_cell(const FieldArgs<T>& a, int i, int j, int k) {
  constexpr bool E = C < 3;
  const int ix[3] = {i, j, k};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (ax == C % 3 || a.s[C][ax] == 1) continue;
    if (E ? ix[ax] == a.s[C][ax] - 1 : ix[ax] == 0) return true;
  }
  return false;
}

template <int C, class T, int N>
__device__ __forceinline__ void load2z(const MainArgs<T>& m, Cells2z<T, N>& q, int i0, int j) {
  constexpr int c = C % 3;
  const FieldArgs<T>& a = m.fa;
  const bool on = ((a.active >> C) & 1u) && j < a.s[C][1];
  const int kc = (a.s[C][2] == 1 || C < 3) ? 0 : 1;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    q.upd[c][n] = false;
    if (!on || i >= a.s[C][0]) continue;
    const int flat = didx(a, C, i, j, kc);
    q.flat[c][n] = flat;
    q.id[c][n] = gld(m.ids[C] + flat);
    q.f[c][n] = gld(a.f[C] + flat);
    if (ghost_cell<C>(a, i, j, kc)) continue;
    q.upd[c][n] = true;
    curl_terms<C>(a, i, j, kc, q.d1[c][n], q.d2[c][n]);
  }
}

template <int C, class T, int N>
__device__ __forceinline__ void update2z(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>* recs,
                                         const Cells2z<T, N>& q, int i0, int j) {
  constexpr int c = C % 3;
  const FieldArgs<T>& a = m.fa;
  const int kc = (a.s[C][2] == 1 || C < 3) ? 0 : 1;
  T* st = static_cast<T*>(m.pml[C]);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    if (!q.upd[c][n]) continue;
    const int flat = q.flat[c][n];
    const T v = store_cell<C>(m, recs[q.id[c][n]], a.f[C], st, flat, i0 + n, j, kc, q.f[c][n], q.d1[c][n],
                              q.d2[c][n]);
    if constexpr (C < 3) {
      if (m.fuse) {  // E ghost planes (the copy k_ghost<E> would make later)
        const int ix[3] = {i0 + n, j, kc};
        fused_ghost<T>(m, C, a.f[C], ix, v);
      }
    }
  }
}

// TMz: the z ghost plane (k = 0) of Hx and Hy, which no E update reads,
// written by the E kernel from k = 1 in place of that part of the H ghost
// copy (k_ghost), which ran just before.  Same values as the copy: Hx's
// corner line with its y ghost plane (j = 0) stays in the copy; Hy's with
// its x ghost plane (i = 0) is phx * (phz * Hy(last, j, 1)), the copy's
// order (z first, then x).
template <class T, int N>
__device__ __forceinline__ void h_zghost(const MainArgs<T>& m, int i0, int j) {
  const FieldArgs<T>& a = m.fa;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    if (((a.active >> 3) & 1u) && a.s[3][2] == 2 && i < a.s[3][0] && j >= 1 && j < a.s[3][1]) {
      const T v = gld(a.f[3] + didx(a, 3, i, j, 1));
      gst(a.f[3] + didx(a, 3, i, j, 0), ghost_phase<T>(m.hzph[0], v));
    }
    if (((a.active >> 4) & 1u) && a.s[4][2] == 2 && i < a.s[4][0] && j < a.s[4][1]) {
      T v = ghost_phase<T>(m.hzph[1], gld(a.f[4] + didx(a, 4, i == 0 ? a.s[4][0] - 1 : i, j, 1)));
      if (i == 0) v = ghost_phase<T>(m.hyxph, v);
      gst(a.f[4] + didx(a, 4, i, j, 0), v);
    }
  }
}

// LDSO: the launch guarantees every block stages its records in LDS, so
// only the LDS copy of the update code is built (see main3d_tile).
template <bool E, class T, int N, bool LDSO = false>
__device__ __forceinline__ void main2z_tile(const MainArgs<T>& m, DevRec<typename real_of<T>::type>* srec, int i0,
                                            int j, int tid, int nthreads) {
  using R = typename real_of<T>::type;
  RecStage<R> stage;
  stage.load(m.recs, m.nrec, tid, nthreads);
  constexpr int C0 = E ? 0 : 3;
  Cells2z<T, N> q;
  load2z<C0 + 0, T, N>(m, q, i0, j);
  load2z<C0 + 1, T, N>(m, q, i0, j);
  load2z<C0 + 2, T, N>(m, q, i0, j);
  stage.commit(srec, m.nrec);
  if (LDSO || stage.use_lds) {
    update2z<C0 + 0, T, N>(m, srec, q, i0, j);
    update2z<C0 + 1, T, N>(m, srec, q, i0, j);
    update2z<C0 + 2, T, N>(m, srec, q, i0, j);
  } else {
    update2z<C0 + 0, T, N>(m, m.recs, q, i0, j);
    update2z<C0 + 1, T, N>(m, m.recs, q, i0, j);
    update2z<C0 + 2, T, N>(m, m.recs, q, i0, j);
  }
  if constexpr (E) {
    if (m.hzg) h_zghost<T, N>(m, i0, j);
  }
}

template <bool E, class T>
__global__ void __launch_bounds__(256) k_main2z(MainArgs<T> m) {
  __shared__ DevRec<typename real_of<T>::type> srec[kRecLds];
  main2z_tile<E, T, k2zRows>(m, srec, blockIdx.y * k2zRows, blockIdx.x * blockDim.x + threadIdx.x, threadIdx.x,
                             blockDim.x);
}

// 3-D main kernel (all six components active).  Thread (i,j,k) owns the
// non-ghost cell of every component of the field type: for E the cells
// (i,j,k) of Ex, Ey, Ez; for H the cells Hx(i,j+1,k+1), Hy(i+1,j,k+1),
// Hz(i+1,j+1,k) -- the ghost planes are never updated, so the thread grid
// is exactly nx*ny*nz.  A block covers a (j,k) tile of N consecutive i
// planes; each curl operand is loaded once and shared by the components
// that use it.  Every field, id and neighbour load of the N planes is
// issued first, through non-aliasing pointers; the record table is staged
// into LDS while those loads are in flight, and only then do the dependent
// record lookups and updates run.
// Real FP32, cells of a component with CPML state: its two psi words are
// loaded with the cell's other operands, so a CPML cell needs no further
// dependent load after its record.  Outside the coordinate runs where
// every CPML profile's c is zero (m.prun) the word is live; inside, it is
// +0 and the lane loads nothing.  A component with UPML state (and no
// CPML) gets its one state word the same way, loaded outside m.ubox, a box
// that holds no UPML cell.
// UP: the launch may carry UPML components (m.upml_mask); kernels without
// it run the shorter CPML-only form (measured 1.2% faster at 256^3).
template <int C, class T, bool UP>
__device__ __forceinline__ void prefetch_psi(const MainArgs<T>& m, int flat, const int ix[3], T& o1, T& o2) {
  o1 = o2 = T{};
  if constexpr (!UP) {
    if (!((m.cpml_mask >> C) & 1u)) return;
    const T* st = static_cast<const T*>(m.pml[C]);
    const int r1 = m.prun[Ax<C>::a1], r2 = m.prun[Ax<C>::a2];
    const int x1 = ix[Ax<C>::a1], x2 = ix[Ax<C>::a2];
    const bool u1 = x1 < (r1 & 0xFFFF) || x1 >= (r1 >> 16);
    const bool u2 = x2 < (r2 & 0xFFFF) || x2 >= (r2 >> 16);
    if (u1) o1 = gld(st + flat);
    if (u2) o2 = gld(st + m.dsize[C] + flat);
    return;
  }
  const bool cp = (m.cpml_mask >> C) & 1u;
  if (!cp && !((m.upml_mask >> C) & 1u)) return;
  const T* st = static_cast<const T*>(m.pml[C]);
  const int r1 = m.prun[Ax<C>::a1], r2 = m.prun[Ax<C>::a2];
  const int x1 = ix[Ax<C>::a1], x2 = ix[Ax<C>::a2];
  bool in = true;  // inside the UPML-free box
#pragma unroll
  for (int a = 0; a < 3; ++a) in = in && ix[a] >= (m.ubox[a] & 0xFFFF) && ix[a] < (m.ubox[a] >> 16);
  const bool u1 = cp ? x1 < (r1 & 0xFFFF) || x1 >= (r1 >> 16) : !in;
  const bool u2 = cp && (x2 < (r2 & 0xFFFF) || x2 >= (r2 >> 16));
  if (u1) o1 = gld(st + flat);
  if (u2) o2 = gld(st + m.dsize[C] + flat);
}

template <class T, int N>
struct Planes {
  T c[7][N];     // curl operands (see load_e / load_h)
  T o[3][N];     // old values of the three updated components
  uint16_t r[3][N];
  T s[3][2][N];  // prefetched CPML psi words (real FP32)
  PRows pr;      // prefetched CPML or UPML profile rows (real FP32, one plane, MainArgs::pp / ur)
};

// The CPML profile rows of coordinates x[0..2] (MainArgs::pp): a lane
// outside the zero-c run of an axis loads that axis's row; inside it the
// row is (b, 0, 1), and b multiplies a psi that is +0 there.
template <class T>
__device__ __forceinline__ void prefetch_rows(const MainArgs<T>& m, const int x[3], PRows& pr) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int r = m.prun[a];
    if (x[a] < (r & 0xFFFF) || x[a] >= (r >> 16)) {
      const GMES_GLOBAL float* p = (const GMES_GLOBAL float*)m.prof + m.pprow[a] + x[a] * 3;
      pr.v[a][0] = p[0];
      pr.v[a][1] = p[1];
      pr.v[a][2] = p[2];
    } else {
      pr.v[a][0] = 0.f;
      pr.v[a][1] = 0.f;
      pr.v[a][2] = 1.f;
    }
  }
}

// The UPML profile rows of a lane outside m.ubox (MainArgs::ur): per axis
// a, c1..c4 of the shared row at xs[a] and c5, c6 of the own row at xo[a].
// Lanes inside the box hold no UPML cell and load nothing.
template <class T>
__device__ __forceinline__ void prefetch_urows(const MainArgs<T>& m, const int xs[3], const int xo[3], bool out,
                                               PRows& pr) {
  if (!out) return;
  const GMES_GLOBAL float* prof = (const GMES_GLOBAL float*)m.prof;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const GMES_GLOBAL float* ps = prof + m.urs[a] + xs[a] * 6;
    const GMES_GLOBAL float* po = prof + m.uro[a] + xo[a] * 6;
    pr.us[a][0] = ps[0];
    pr.us[a][1] = ps[1];
    pr.us[a][2] = ps[2];
    pr.us[a][3] = ps[3];
    pr.uo[a][0] = po[4];
    pr.uo[a][1] = po[5];
  }
}

// Whether storage coordinates x lie outside the UPML-free box m.ubox.
template <class T>
__device__ __forceinline__ bool outside_ubox(const MainArgs<T>& m, const int x[3]) {
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) in = in && x[a] >= (m.ubox[a] & 0xFFFF) && x[a] < (m.ubox[a] >> 16);
  return !in;
}

template <class T, int N, bool UP, bool PP = false, bool UR = false>
__device__ __forceinline__ void load_e(const MainArgs<T>& m, Planes<T, N>& q, int i0, int j, int k, const int* B, const int* P,
                                       const int* W, const T* __restrict__ hx, const T* __restrict__ hy,
                                       const T* __restrict__ hz, const T* __restrict__ ex, const T* __restrict__ ey,
                                       const T* __restrict__ ez, const uint16_t* __restrict__ idx,
                                       const uint16_t* __restrict__ idy, const uint16_t* __restrict__ idz, const uint8_t* __restrict__ bx8,
                                       const uint8_t* __restrict__ by8, const uint8_t* __restrict__ bz8,
                                       bool narrow) {
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    q.c[0][n] = gld(hx + B[3] + i * P[3] + (j + 1) * W[3] + k + 1);  // hx(i, j+1, k+1)
    q.c[1][n] = gld(hx + B[3] + i * P[3] + j * W[3] + k + 1);        // hx(i, j,   k+1)
    q.c[2][n] = gld(hx + B[3] + i * P[3] + (j + 1) * W[3] + k);      // hx(i, j+1, k)
    q.c[3][n] = gld(hy + B[4] + (i + 1) * P[4] + j * W[4] + k + 1);  // hy(i+1, j, k+1)
    q.c[4][n] = gld(hy + B[4] + (i + 1) * P[4] + j * W[4] + k);      // hy(i+1, j, k)
    q.c[5][n] = gld(hz + B[5] + (i + 1) * P[5] + (j + 1) * W[5] + k);  // hz(i+1, j+1, k)
    q.c[6][n] = gld(hz + B[5] + (i + 1) * P[5] + j * W[5] + k);        // hz(i+1, j,   k)
    const int fx = B[0] + i * P[0] + j * W[0] + k, fy = B[1] + i * P[1] + j * W[1] + k;
    const int fz = B[2] + i * P[2] + j * W[2] + k;
    q.o[0][n] = ld_nt(ex + fx);
    q.o[1][n] = ld_nt(ey + fy);
    q.o[2][n] = ld_nt(ez + fz);
    if (sizeof(T) == 4 && narrow) {  // byte ids: real FP32 only (measured)
      q.r[0][n] = ld_nt(bx8 + fx);
      q.r[1][n] = ld_nt(by8 + fy);
      q.r[2][n] = ld_nt(bz8 + fz);
    } else {
      q.r[0][n] = ld_nt(idx + fx);
      q.r[1][n] = ld_nt(idy + fy);
      q.r[2][n] = ld_nt(idz + fz);
    }
    if constexpr (sizeof(T) == 4) {
      const int ix[3] = {i, j, k};
      prefetch_psi<0, T, UP>(m, fx, ix, q.s[0][0][n], q.s[0][1][n]);
      prefetch_psi<1, T, UP>(m, fy, ix, q.s[1][0][n], q.s[1][1][n]);
      prefetch_psi<2, T, UP>(m, fz, ix, q.s[2][0][n], q.s[2][1][n]);
      if constexpr (PP) prefetch_rows<T>(m, ix, q.pr);  // E: x at i, y at j, z at k
      if constexpr (UR) prefetch_urows<T>(m, ix, ix, outside_ubox(m, ix), q.pr);  // E: all at (i, j, k)
    }
  }
}

// prev: H'z(i0-1,j,k) = Hz[i0, j+1, k] and H'y(i0-1,j,k) = Hy[i0, j, k+1],
// the i-1 operands of the first plane; advanced to the last plane's.
template <class T, int N, bool PP = false, bool UR = false>
__device__ __forceinline__ void update_e(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>* recs,
                                         const Planes<T, N>& q, int i0, int j, int k, const int* B, const int* P,
                                         const int* W, T* __restrict__ ex, T* __restrict__ ey, T* __restrict__ ez,
                                         T* __restrict__ sx, T* __restrict__ sy, T* __restrict__ sz, T& hz_prev,
                                         T& hy_prev) {
  constexpr bool kPre = sizeof(T) == 4;  // psi words prefetched (load_e / load_h)
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    const int fx = B[0] + i * P[0] + j * W[0] + k, fy = B[1] + i * P[1] + j * W[1] + k;
    const int fz = B[2] + i * P[2] + j * W[2] + k;
    const T vx = store_cell<0, T, kPre, PP, UR>(m, recs[q.r[0][n]], ex, sx, fx, i, j, k, q.o[0][n],
                                            vsub(q.c[5][n], q.c[6][n]), vsub(q.c[3][n], q.c[4][n]),
                                            q.s[0][0][n], q.s[0][1][n], q.pr);
    const T vy = store_cell<1, T, kPre, PP, UR>(m, recs[q.r[1][n]], ey, sy, fy, i, j, k, q.o[1][n],
                                            vsub(q.c[0][n], q.c[2][n]), vsub(q.c[5][n], hz_prev),
                                            q.s[1][0][n], q.s[1][1][n], q.pr);
    const T vz = store_cell<2, T, kPre, PP, UR>(m, recs[q.r[2][n]], ez, sz, fz, i, j, k, q.o[2][n],
                                            vsub(q.c[3][n], hy_prev), vsub(q.c[0][n], q.c[1][n]),
                                            q.s[2][0][n], q.s[2][1][n], q.pr);
    if (m.fuse && (i == 0 || j == 0 || k == 0)) {
      const int ix[3] = {i, j, k};
      fused_ghost<T>(m, 0, ex, ix, vx);
      fused_ghost<T>(m, 1, ey, ix, vy);
      fused_ghost<T>(m, 2, ez, ix, vz);
    }
    hz_prev = q.c[5][n];
    hy_prev = q.c[3][n];
  }
}

template <class T, int N, bool UP, bool PP = false, bool UR = false>
__device__ __forceinline__ void load_h(const MainArgs<T>& m, Planes<T, N>& q, int i0, int j, int k, const int* B, const int* P,
                                       const int* W, const T* __restrict__ ex, const T* __restrict__ ey,
                                       const T* __restrict__ ez, const T* __restrict__ hx, const T* __restrict__ hy,
                                       const T* __restrict__ hz, const uint16_t* __restrict__ idx,
                                       const uint16_t* __restrict__ idy, const uint16_t* __restrict__ idz, const uint8_t* __restrict__ bx8,
                                       const uint8_t* __restrict__ by8, const uint8_t* __restrict__ bz8,
                                       bool narrow) {
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    q.c[0][n] = gld(ex + B[0] + i * P[0] + j * W[0] + k);            // ex(i, j, k)
    q.c[1][n] = gld(ex + B[0] + i * P[0] + j * W[0] + k + 1);        // ex(i, j, k+1)
    q.c[2][n] = gld(ex + B[0] + i * P[0] + (j + 1) * W[0] + k);      // ex(i, j+1, k)
    q.c[3][n] = gld(ey + B[1] + i * P[1] + j * W[1] + k + 1);        // ey(i, j, k+1)
    q.c[4][n] = gld(ey + B[1] + (i + 1) * P[1] + j * W[1] + k);      // ey(i+1, j, k)
    q.c[5][n] = gld(ez + B[2] + i * P[2] + (j + 1) * W[2] + k);      // ez(i, j+1, k)
    q.c[6][n] = gld(ez + B[2] + (i + 1) * P[2] + j * W[2] + k);      // ez(i+1, j, k)
    const int fx = B[3] + i * P[3] + (j + 1) * W[3] + k + 1;    // Hx(i, j+1, k+1)
    const int fy = B[4] + (i + 1) * P[4] + j * W[4] + k + 1;    // Hy(i+1, j, k+1)
    const int fz = B[5] + (i + 1) * P[5] + (j + 1) * W[5] + k;  // Hz(i+1, j+1, k)
    q.o[0][n] = ld_nt(hx + fx);
    q.o[1][n] = ld_nt(hy + fy);
    q.o[2][n] = ld_nt(hz + fz);
    if (sizeof(T) == 4 && narrow) {  // byte ids: real FP32 only (measured)
      q.r[0][n] = ld_nt(bx8 + fx);
      q.r[1][n] = ld_nt(by8 + fy);
      q.r[2][n] = ld_nt(bz8 + fz);
    } else {
      q.r[0][n] = ld_nt(idx + fx);
      q.r[1][n] = ld_nt(idy + fy);
      q.r[2][n] = ld_nt(idz + fz);
    }
    if constexpr (sizeof(T) == 4) {
      const int ixx[3] = {i, j + 1, k + 1}, ixy[3] = {i + 1, j, k + 1}, ixz[3] = {i + 1, j + 1, k};
      prefetch_psi<3, T, UP>(m, fx, ixx, q.s[0][0][n], q.s[0][1][n]);
      prefetch_psi<4, T, UP>(m, fy, ixy, q.s[1][0][n], q.s[1][1][n]);
      prefetch_psi<5, T, UP>(m, fz, ixz, q.s[2][0][n], q.s[2][1][n]);
      if constexpr (PP) {
        const int x[3] = {i + 1, j + 1, k + 1};  // H: x at i+1, y at j+1, z at k+1
        prefetch_rows<T>(m, x, q.pr);
      }
      if constexpr (UR) {
        // shared rows at i+1, j+1, k+1; own rows (Hx along x, Hy along y,
        // Hz along z) at i, j, k
        const int xs[3] = {i + 1, j + 1, k + 1}, xo[3] = {i, j, k};
        prefetch_urows<T>(m, xs, xo, outside_ubox(m, ixx) || outside_ubox(m, ixy) || outside_ubox(m, ixz), q.pr);
      }
    }
  }
}

// cur: Ey(i0,j,k), Ez(i0,j,k), advanced to the last plane's i+1 values.
template <class T, int N, bool PP = false, bool UR = false>
__device__ __forceinline__ void update_h(const MainArgs<T>& m, const DevRec<typename real_of<T>::type>* recs,
                                         const Planes<T, N>& q, int i0, int j, int k, const int* B, const int* P,
                                         const int* W, T* __restrict__ hx, T* __restrict__ hy, T* __restrict__ hz,
                                         T* __restrict__ sx, T* __restrict__ sy, T* __restrict__ sz, T& ey_cur,
                                         T& ez_cur) {
  constexpr bool kPre = sizeof(T) == 4;  // psi words prefetched (load_e / load_h)
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int i = i0 + n;
    const int fx = B[3] + i * P[3] + (j + 1) * W[3] + k + 1;
    const int fy = B[4] + (i + 1) * P[4] + j * W[4] + k + 1;
    const int fz = B[5] + (i + 1) * P[5] + (j + 1) * W[5] + k;
    const T vx = store_cell<3, T, kPre, PP, UR>(m, recs[q.r[0][n]], hx, sx, fx, i, j + 1, k + 1, q.o[0][n],
                                            vsub(q.c[5][n], ez_cur), vsub(q.c[3][n], ey_cur), q.s[0][0][n],
                                            q.s[0][1][n], q.pr);
    const T vy = store_cell<4, T, kPre, PP, UR>(m, recs[q.r[1][n]], hy, sy, fy, i + 1, j, k + 1, q.o[1][n],
                                            vsub(q.c[1][n], q.c[0][n]), vsub(q.c[6][n], ez_cur),
                                            q.s[1][0][n], q.s[1][1][n], q.pr);
    const T vz = store_cell<5, T, kPre, PP, UR>(m, recs[q.r[2][n]], hz, sz, fz, i + 1, j + 1, k, q.o[2][n],
                                            vsub(q.c[4][n], ey_cur), vsub(q.c[2][n], q.c[0][n]),
                                            q.s[2][0][n], q.s[2][1][n], q.pr);
    ey_cur = q.c[4][n];
    ez_cur = q.c[6][n];
  }
}


// One tile of the 3-D update: thread (k, j) of planes i0 .. i0+N-1; tid and
// nthreads are the thread's rank in the block and the block size (the
// record-table staging and its barrier involve every thread of the block).
// SPLIT: separate update code for records in LDS and in global memory
// (faster for the single-sim kernel); otherwise one copy reads them
// through a generic pointer (faster, as measured, for the batched kernel).
// LDSO: the launch guarantees that every block stages its records in LDS
// (RecStage::use_lds), so one copy reads them from LDS only (batched
// kernel, 7% faster than the generic pointer).  UP: see prefetch_psi.
template <bool E, class T, int N, bool SPLIT, bool UP, bool LDSO = false, bool PP = false, bool UR = false>
__device__ __forceinline__ void main3d_tile(const MainArgs<T>& m, DevRec<typename real_of<T>::type>* srec, int k,
                                            int j, int i0, int tid, int nthreads) {
  static_assert(!PP || (N == 1 && LDSO && std::is_same<T, float>::value), "row prefetch: real FP32, one plane, LDS records");
  static_assert(!UR || (N == 1 && LDSO && UP && !PP && std::is_same<T, float>::value),
                "UPML row prefetch: real FP32, one plane, LDS records, UPML state prefetch");
  using R = typename real_of<T>::type;
  const int nx = m.n[0], ny = m.n[1], nz = m.n[2];
  const bool act = k < nz && j < ny;  // every thread reaches the barrier
  const FieldArgs<T>& a = m.fa;
  // per component: element of (i,j,k) = B + i*P + j*W + k (pitched layout)
  int B[6], P[6], W[6];
  for (int c = 0; c < 6; ++c) { B[c] = kPad - a.koff[c]; W[c] = a.p[c]; P[c] = a.s[c][1] * a.p[c]; }
  T* const* f = a.f;
  T* sx = static_cast<T*>(m.pml[E ? 0 : 3]);
  T* sy = static_cast<T*>(m.pml[E ? 1 : 4]);
  T* sz = static_cast<T*>(m.pml[E ? 2 : 5]);
  const uint16_t* ida = m.ids[E ? 0 : 3];
  const uint16_t* idb = m.ids[E ? 1 : 4];
  const uint16_t* idc = m.ids[E ? 2 : 5];
  const uint8_t* i8a = m.ids8[E ? 0 : 3];
  const uint8_t* i8b = m.ids8[E ? 1 : 4];
  const uint8_t* i8c = m.ids8[E ? 2 : 5];
  const DevRec<R>* grec = m.recs;
  RecStage<R> stage;
  stage.load(grec, m.nrec, tid, nthreads);
  if (i0 + N <= nx) {
    Planes<T, N> q;
    T u0 = T{}, u1 = T{};
    if (act) {
      if constexpr (E) {
        load_e<T, N, UP, PP, UR>(m, q, i0, j, k, B, P, W, f[3], f[4], f[5], f[0], f[1], f[2], ida, idb, idc, i8a, i8b, i8c, m.narrow != 0);
        u0 = gld(f[5] + B[5] + i0 * P[5] + (j + 1) * W[5] + k);
        u1 = gld(f[4] + B[4] + i0 * P[4] + j * W[4] + k + 1);
      } else {
        load_h<T, N, UP, PP, UR>(m, q, i0, j, k, B, P, W, f[0], f[1], f[2], f[3], f[4], f[5], ida, idb, idc, i8a, i8b, i8c, m.narrow != 0);
        u0 = gld(f[1] + B[1] + i0 * P[1] + j * W[1] + k);
        u1 = gld(f[2] + B[2] + i0 * P[2] + j * W[2] + k);
      }
    }
    stage.commit(srec, m.nrec);
    if (!act) return;
    if constexpr (LDSO) {
      if constexpr (E) update_e<T, N, PP, UR>(m, srec, q, i0, j, k, B, P, W, f[0], f[1], f[2], sx, sy, sz, u0, u1);
      else update_h<T, N, PP, UR>(m, srec, q, i0, j, k, B, P, W, f[3], f[4], f[5], sx, sy, sz, u0, u1);
    } else if constexpr (!SPLIT) {
      const DevRec<R>* recs = stage.use_lds ? srec : grec;
      if constexpr (E) update_e<T, N>(m, recs, q, i0, j, k, B, P, W, f[0], f[1], f[2], sx, sy, sz, u0, u1);
      else update_h<T, N>(m, recs, q, i0, j, k, B, P, W, f[3], f[4], f[5], sx, sy, sz, u0, u1);
    } else if (stage.use_lds) {
      if constexpr (E) update_e<T, N>(m, srec, q, i0, j, k, B, P, W, f[0], f[1], f[2], sx, sy, sz, u0, u1);
      else update_h<T, N>(m, srec, q, i0, j, k, B, P, W, f[3], f[4], f[5], sx, sy, sz, u0, u1);
    } else {
      if constexpr (E) update_e<T, N>(m, grec, q, i0, j, k, B, P, W, f[0], f[1], f[2], sx, sy, sz, u0, u1);
      else update_h<T, N>(m, grec, q, i0, j, k, B, P, W, f[3], f[4], f[5], sx, sy, sz, u0, u1);
    }
  } else {
    stage.commit(srec, m.nrec);
    if (!act) return;
    const DevRec<R>* recs = LDSO ? srec : (stage.use_lds ? srec : grec);  // remainder planes only
    T u0, u1;
    if constexpr (E) {
      u0 = gld(f[5] + B[5] + i0 * P[5] + (j + 1) * W[5] + k);
      u1 = gld(f[4] + B[4] + i0 * P[4] + j * W[4] + k + 1);
    } else {
      u0 = gld(f[1] + B[1] + i0 * P[1] + j * W[1] + k);
      u1 = gld(f[2] + B[2] + i0 * P[2] + j * W[2] + k);
    }
    for (int i = i0; i < nx; ++i) {
      Planes<T, 1> q;
      if constexpr (E) {
        load_e<T, 1, UP>(m, q, i, j, k, B, P, W, f[3], f[4], f[5], f[0], f[1], f[2], ida, idb, idc, i8a, i8b, i8c, m.narrow != 0);
        update_e<T, 1>(m, recs, q, i, j, k, B, P, W, f[0], f[1], f[2], sx, sy, sz, u0, u1);
      } else {
        load_h<T, 1, UP>(m, q, i, j, k, B, P, W, f[0], f[1], f[2], f[3], f[4], f[5], ida, idb, idc, i8a, i8b, i8c, m.narrow != 0);
        update_h<T, 1>(m, recs, q, i, j, k, B, P, W, f[3], f[4], f[5], sx, sy, sz, u0, u1);
      }
    }
  }
}

// Waves per SIMD the register allocator must leave room for (L: the
// LDS-only record variant, which needs far fewer registers).  Real FP32
// with one plane per thread fits 64 VGPRs (8 waves); with split record
// copies it gets there by moving a few scalar registers to VGPR lanes,
// 0.8% faster on 256^3 than 7 waves.  The LDS-only FP32 kernels reach 8
// waves without the hint too, but ran 1% slower without it.  FP64 with split copies forced from 5
// to 6 waves spills to scratch and is 13% slower, so it stays free; the
// LDS-only FP64 kernel fits 7 waves without spills (256^3 0.784 to 0.730
// ms per step), while 8 waves spills and is 4% slower than 7.
template <class T, int N, bool L = false> struct MainWaves { static constexpr int v = 1; };
template <bool L> struct MainWaves<float, 1, L> { static constexpr int v = 8; };
template <> struct MainWaves<double, 1, true> { static constexpr int v = 7; };

// LDSO: the launch guarantees the record table fits the LDS staging (see
// main3d_tile); otherwise the update code is split into an LDS and a
// global copy.
template <bool E, class T, int N, bool UP, bool LDSO, bool PP = false, bool UR = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MainWaves<T, N, LDSO>::v)))
k_main3d(MainArgs<T> m) {
  __shared__ DevRec<typename real_of<T>::type> srec[kRecLds];
  main3d_tile<E, T, N, !LDSO, UP, LDSO, PP, UR>(m, srec, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                       blockIdx.z * N, threadIdx.y * blockDim.x + threadIdx.x, blockDim.x * blockDim.y);
}

// UR in the flat and batched kernels: the UPML rows need more registers
// than 8 waves per SIMD allow (64 VGPRs: 26 to 40 spilled); at 7 waves the
// kernels fit in 66.
constexpr int kUrWaves = 7;

// Flat single-sim launch (packed rows, gmes_sim::packed): block b takes the
// 256 consecutive row-major (j, k) points b % gx of plane b / gx, with
// gx = ceil(ny * nz / 256), so short rows leave no lane idle.
template <bool E, class T, bool UP, bool LDSO, bool PP = false, bool UR = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UR ? kUrWaves : MainWaves<T, 1, LDSO>::v)))
k_main3d_flat(MainArgs<T> m) {
  __shared__ DevRec<typename real_of<T>::type> srec[kRecLds];
  const unsigned nz = (unsigned)m.n[2];
  const unsigned gx = ((unsigned)m.n[1] * nz + 255u) / 256u;
  const unsigned i = blockIdx.x / gx;
  const unsigned f = (blockIdx.x - i * gx) * 256u + threadIdx.x;
  const unsigned j = f / nz;
  main3d_tile<E, T, 1, !LDSO, UP, LDSO, PP, UR>(m, srec, (int)(f - j * nz), (int)j, (int)i, threadIdx.x, 256);
}

// ------------------------------------------------------------ batched launches
// Many sims of one storage type advance with one launch per kernel kind:
// a job table in device memory holds each member's arguments and the index
// of its first block; a block finds its job by binary search on that index
// (uniform across the block) and runs the same tile body as the single-sim
// kernel (MainJob: which points a block takes).

template <bool E, class T, bool LDSO, bool PP = false, bool UR = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UR ? kUrWaves : MainWaves<T, 1>::v)))
k_main3d_batch(const MainJob<T>* jobs, int njobs) {
  __shared__ DevRec<typename real_of<T>::type> srec[kRecLds];
  const MainJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const unsigned local = (unsigned)(blockIdx.x - J.block0);
  const unsigned i = local / (unsigned)J.gx;
  const unsigned f = (local - i * (unsigned)J.gx) * 256u + threadIdx.x;  // row-major (j, k) of plane i
  const unsigned j = f / (unsigned)J.m.n[2];
  // PP members have CPML state in every component of the field type, so no
  // UPML state word to prefetch (UP = false, the shorter form)
  main3d_tile<E, T, 1, false, !PP, LDSO, PP, UR>(J.m, srec, (int)(f - j * (unsigned)J.m.n[2]), (int)j, (int)i,
                                              threadIdx.x, 256);
}

template <bool E, class T>
__global__ void __launch_bounds__(256) k_main_batch(const MainJob<T>* jobs, int njobs) {
  const MainJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const int local = (int)(blockIdx.x - J.block0);
  const int bxi = local % J.gx, byi = (local / J.gx) % J.gy, bzi = local / (J.gx * J.gy);
  const int tid = threadIdx.x;
  const int tx = tid & (J.bx - 1), ty = tid / J.bx;
  const int j = byi * (256 / J.bx) + ty;
  if (J.m.kloop) {
    for (int k = 0; k < J.m.kext; ++k) main_generic_cell<E, T>(J.m, bzi, j, k);
  } else {
    main_generic_cell<E, T>(J.m, bzi, j, bxi * J.bx + tx);
  }
}

// Collapsed-z 2-D members: a job's gx blocks each take 256 consecutive
// (i, j) columns of the row-major i x bx plane (bx: the extent along j).
template <bool E, class T, bool LDSO>
__global__ void __launch_bounds__(256) k_main2z_batch(const MainJob<T>* jobs, int njobs) {
  __shared__ DevRec<typename real_of<T>::type> srec[kRecLds];
  const MainJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const unsigned f = (unsigned)(blockIdx.x - J.block0) * 256u + threadIdx.x;
  const unsigned i = f / (unsigned)J.bx;
  main2z_tile<E, T, 1, LDSO>(J.m, srec, (int)i, (int)(f - i * (unsigned)J.bx), threadIdx.x, 256);
}

// Batched launches: block -> job by binary search on the jobs' first
// blocks (uniform across a block).
template <class J>
__device__ __forceinline__ int find_job(const J* jobs, int njobs, int64_t b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ----------------------------------------------------------------- ADE kernels
template <class T>
__device__ __forceinline__ void cell_ijk(const FieldArgs<T>& a, int c, int flat, int& i, int& j, int& k) {
  const int sz = a.s[c][2], sy = a.s[c][1];
  k = flat % sz;
  const int t = flat / sz;
  j = t % sy;
  i = t / sy;
}

// Drude and Lorentz poles: src/pw_drude.hh:90-101,145-150; src/pw_lorentz.hh:96-107,150-156
template <int C, class T>
__device__ __forceinline__ void ade_drude_cell(const AdeArgs<T>& m, int q) {
  using R = typename real_of<T>::type;
  if (q >= m.n) return;
  const int flat = m.cell[q];
  int i, j, k;
  cell_ijk(m.fa, C, flat, i, j, k);
  const DevRec<R>& rec = m.recs[m.rec[q]];
  const R* a = m.coef + rec.coef_off;
  const int P = rec.n1;
  const R* cc = a + 3 * P;
  T* st = static_cast<T*>(m.state);
  T* F = m.fa.f[C] + didx(m.fa, C, i, j, k);
  const T e_now = ld_nt(F);
  T sum = vzero<T>();
  for (int p = 0; p < P; ++p) {
    const T q_old = ld_nt(st + p * m.n + q);
    const T q_now = ld_nt(st + (P + p) * m.n + q);
    const T q_new = vadd(vadd(vscl(a[3 * p], q_old), vscl(a[3 * p + 1], q_now)), vscl(a[3 * p + 2], e_now));
    st_nt(st + p * m.n + q, q_now);
    st_nt(st + (P + p) * m.n + q, q_new);
    sum = vadd(sum, vsub(q_new, q_now));
  }
  T D1, D2;
  curl_terms<C>(m.fa, i, j, k, D1, D2);
  const R d1 = m.d[Ax<C>::a1], d2 = m.d[Ax<C>::a2];
  T cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
  st_nt(F, vadd(vadd(vscl(cc[0], cu), vscl(cc[1], sum)), vscl(cc[2], e_now)));
}

template <int C, class T>
__global__ void __launch_bounds__(256) k_ade_drude(AdeArgs<T> m) {
  ade_drude_cell<C, T>(m, blockIdx.x * blockDim.x + threadIdx.x);
}

// Drude-critical-points, ADE form: src/pw_dcp.hh:105-167,200-225
template <int C, class T>
__device__ __forceinline__ void ade_dcp_cell(const AdeArgs<T>& m, int q) {
  using R = typename real_of<T>::type;
  if (q >= m.n) return;
  const int flat = m.cell[q];
  int i, j, k;
  cell_ijk(m.fa, C, flat, i, j, k);
  const DevRec<R>& rec = m.recs[m.rec[q]];
  const int D = rec.n1, NC = rec.n2;
  const R* a = m.coef + rec.coef_off;
  const R* b = a + 3 * D;
  const R* cc = b + 5 * NC;
  T* st = static_cast<T*>(m.state);
  const int n = m.n;
  // words: 0 e_old | 1..D q_old | D+1..2D q_now | 2D+1.. p_old | 2D+C+1.. p_now
  T* F = m.fa.f[C] + didx(m.fa, C, i, j, k);
  const T e_now = *F;
  const T e_old = st[q];
  T ds = vzero<T>(), cs = vzero<T>();
  for (int p = 0; p < D; ++p)
    ds = vadd(ds, vsub(vscl(R(1) - a[3 * p + 1], st[(1 + D + p) * n + q]), vscl(a[3 * p], st[(1 + p) * n + q])));
  for (int p = 0; p < NC; ++p)
    cs = vadd(cs, vsub(vscl(R(1) - b[5 * p + 1], st[(1 + 2 * D + NC + p) * n + q]),
                       vscl(b[5 * p], st[(1 + 2 * D + p) * n + q])));
  T D1, D2;
  curl_terms<C>(m.fa, i, j, k, D1, D2);
  const R d1 = m.d[Ax<C>::a1], d2 = m.d[Ax<C>::a2];
  T cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
  T e_new = vadd(vadd(vadd(vscl(cc[0], cu), vscl(cc[1], vadd(ds, cs))), vscl(cc[2], e_old)), vscl(cc[3], e_now));
  T mix = vadd(vadd(e_old, vscl(R(2), e_now)), e_new);
  for (int p = 0; p < D; ++p) {
    T qo = st[(1 + p) * n + q], qn = st[(1 + D + p) * n + q];
    T qq = vadd(vadd(vscl(a[3 * p], qo), vscl(a[3 * p + 1], qn)), vscl(a[3 * p + 2], mix));
    st_nt(st + (1 + p) * n + q, qn);
    st_nt(st + (1 + D + p) * n + q, qq);
  }
  for (int p = 0; p < NC; ++p) {
    T po = st[(1 + 2 * D + p) * n + q], pn = st[(1 + 2 * D + NC + p) * n + q];
    T pp = vadd(vadd(vadd(vadd(vscl(b[5 * p], po), vscl(b[5 * p + 1], pn)), vscl(b[5 * p + 2], e_old)),
                     vscl(b[5 * p + 3], e_now)),
                vscl(b[5 * p + 4], e_new));
    st_nt(st + (1 + 2 * D + p) * n + q, pn);
    st_nt(st + (1 + 2 * D + NC + p) * n + q, pp);
  }
  st_nt(st + q, e_now);
  st_nt(F, e_new);
}

template <int C, class T>
__global__ void __launch_bounds__(256) k_ade_dcp(AdeArgs<T> m) {
  ade_dcp_cell<C, T>(m, blockIdx.x * blockDim.x + threadIdx.x);
}

// Drude-critical-points, PLRC/RC form: src/pw_dcp.hh:467-518,564-586.
// State words are complex for any field type: psi_dp (re part = psi_dp_re,
// im part = psi_dp_im), psi_cp_re, psi_cp_im.
template <int C, class T>
__device__ __forceinline__ void ade_plrc_cell(const AdeArgs<T>& m, int q) {
  using R = typename real_of<T>::type;
  using Z = cplx<R>;
  if (q >= m.n) return;
  const int flat = m.cell[q];
  int i, j, k;
  cell_ijk(m.fa, C, flat, i, j, k);
  const DevRec<R>& rec = m.recs[m.rec[q]];
  const int D = rec.n1, NC = rec.n2;
  const R* a = m.coef + rec.coef_off;
  const R* b = a + 3 * D;
  const R* cc = b + 6 * NC;
  Z* st = static_cast<Z*>(m.state);
  const int n = m.n;
  T* F = m.fa.f[C] + didx(m.fa, C, i, j, k);
  const T fv = *F;
  const Z e_now = {re_of(fv), im_of(fv)};
  R sdr = R(0), sdi = R(0);
  for (int p = 0; p < D; ++p) sdr = sdr + st[p * n + q].re;
  for (int p = 0; p < D; ++p) sdi = sdi + st[p * n + q].im;
  Z scr = {R(0), R(0)}, sci = {R(0), R(0)};
  for (int p = 0; p < NC; ++p) scr = vadd(scr, st[(D + p) * n + q]);
  for (int p = 0; p < NC; ++p) sci = vadd(sci, st[(D + NC + p) * n + q]);
  const Z psi = {sdr + scr.re, sdi + sci.re};  // .real() of the imaginary-part sum
  T D1, D2;
  curl_terms<C>(m.fa, i, j, k, D1, D2);
  const R d1 = m.d[Ax<C>::a1], d2 = m.d[Ax<C>::a2];
  const T cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
  Z e_new;
  if constexpr (is_cplx<T>::value) {
    e_new = vadd(vadd(vscl(cc[0], Z{cu.re, cu.im}), vscl(cc[1], e_now)), vscl(cc[2], psi));
  } else {
    e_new.re = (cc[0] * cu + cc[1] * e_now.re) + cc[2] * psi.re;
    e_new.im = cc[1] * e_now.im + cc[2] * psi.im;
  }
  for (int p = 0; p < D; ++p) {
    Z v = st[p * n + q];
    v.re = (a[3 * p] * e_new.re + a[3 * p + 1] * e_now.re) + a[3 * p + 2] * v.re;
    v.im = (a[3 * p] * e_new.im + a[3 * p + 1] * e_now.im) + a[3 * p + 2] * v.im;
    st[p * n + q] = v;
  }
  for (int p = 0; p < NC; ++p) {
    const Z b0 = {b[6 * p], b[6 * p + 1]}, b1 = {b[6 * p + 2], b[6 * p + 3]}, b2 = {b[6 * p + 4], b[6 * p + 5]};
    st[(D + p) * n + q] = vadd(vadd(vscl(e_new.re, b0), vscl(e_now.re, b1)), cmul(b2, st[(D + p) * n + q]));
    st[(D + NC + p) * n + q] = vadd(vadd(vscl(e_new.im, b0), vscl(e_now.im, b1)), cmul(b2, st[(D + NC + p) * n + q]));
  }
  *F = make_val<T>(e_new.re, e_new.im);
}

template <int C, class T>
__global__ void __launch_bounds__(256) k_ade_plrc(AdeArgs<T> m) {
  ade_plrc_cell<C, T>(m, blockIdx.x * blockDim.x + threadIdx.x);
}

// Two-level density matrix, predictor-corrector: src/pw_dm2.hh:345-407.
// Real fields only; uses only the D2 curl term and no eps (reference quirk).
template <class R>
__device__ double dm2_l2(R e, const R* u, int n) {
  R accu = e * e;
  for (int i = 0; i < n; ++i) {
    R ip = R(0);
    for (int jj = 0; jj < 3; ++jj) ip = ip + u[3 * i + jj] * u[3 * i + jj];
    accu += ip;
  }
  return sqrt((double)accu);
}

template <int C, class T>
__device__ __forceinline__ void ade_dm2_cell(const AdeArgs<T>& m, int q) {
  using R = typename real_of<T>::type;
  if constexpr (is_cplx<T>::value) {
    return;
  } else {
      if (q >= m.n) return;
    const int flat = m.cell[q];
    int i, j, k;
    cell_ijk(m.fa, C, flat, i, j, k);
    const DevRec<R>& rec = m.recs[m.rec[q]];
    const int N = rec.n1;
    const R* cf = m.coef + rec.coef_off;
    const R* omega = cf;
    const R* n_atom = cf + N;
    const R rho30 = cf[2 * N], gamma = cf[2 * N + 1], t1 = cf[2 * N + 2], t2 = cf[2 * N + 3];
    const R hbar = cf[2 * N + 4];
    const double rtol = (double)cf[2 * N + 5];
    const int64_t step = m.ctl->clock - 1;
    const R n_half = (R)((double)step + 0.5);
    const R t = (n_half + R(0.5)) * m.dt;
    R av[kDm2Max], bv[kDm2Max], u[3 * kDm2Max], u_new[3 * kDm2Max], u_tmp[3 * kDm2Max], dif[3 * kDm2Max];
    for (int p = 0; p < N; ++p) {
      av[p] = n_atom[p] * gamma / t2 * exp(-t / t2);
      bv[p] = n_atom[p] * gamma * omega[p] * exp(-t / t2);
    }
    const R c_plus = R(2) / hbar * gamma * exp(-t * (R(1) / t1 - R(1) / t2));
    const R c_minus = R(2) / hbar * gamma * exp(-t * (R(1) / t2 - R(1) / t1));
    const R dd = R(2) / hbar * gamma * rho30 * exp(t / t2);
    R* st = static_cast<R*>(m.state);
    for (int p = 0; p < 3 * N; ++p) { u[p] = st[p * m.n + q]; u_new[p] = u[p]; }
    T* F = m.fa.f[C] + didx(m.fa, C, i, j, k);
    R e_new = *F;
    const R e_old = *F;
    T D1, D2;
    curl_terms<C>(m.fa, i, j, k, D1, D2);
    const R h_d = D2 / m.d[Ax<C>::a2];
    const R dt = m.dt;
    double error;
    int iteration = 0, fail = 0;
    do {
      const R e_tmp = e_new;
      for (int p = 0; p < 3 * N; ++p) u_tmp[p] = u_new[p];
      e_new = e_old - dt * h_d;
      for (int p = 0; p < N; ++p) {
        e_new -= R(.5) * dt * av[p] * (u_new[3 * p] + u[3 * p]);
        e_new += R(.5) * dt * bv[p] * (u_new[3 * p + 1] + u[3 * p + 1]);
      }
      for (int p = 0; p < N; ++p) {
        u_new[3 * p] = u[3 * p] + R(.5) * dt * omega[p] * (u_new[3 * p + 1] + u[3 * p + 1]);
        u_new[3 * p + 1] = u[3 * p + 1] - R(.5) * dt * omega[p] * (u_new[3 * p] + u[3 * p]) +
                           R(.25) * dt * c_plus * (u_new[3 * p + 2] + u[3 * p + 2]) * (e_new + e_old) +
                           R(.5) * dt * dd * (e_new + e_old);
        u_new[3 * p + 2] = u[3 * p + 2] - R(.25) * dt * c_minus * (u_new[3 * p + 1] + u[3 * p + 1]) * (e_new + e_old);
      }
      for (int p = 0; p < 3 * N; ++p) dif[p] = u_new[p] - u_tmp[p];
      const double dn = dm2_l2(e_new - e_tmp, dif, N);
      const double rn = dm2_l2(e_tmp, u_tmp, N);
      error = rn == 0 ? (dn == 0 ? 0.0 : INFINITY) : dn / rn;
      if (isnan(error)) { fail = 1; break; }
      if (++iteration >= 100 && error > rtol) { fail = 2; break; }
    } while (error > rtol);
    if (fail) {
      m.ctl->err = fail;  // plain vector store; any failing cell reports
      return;
    }
    *F = e_new;
    for (int p = 0; p < 3 * N; ++p) st[p * m.n + q] = u_new[p];
  }
}

template <int C, class T>
__global__ void __launch_bounds__(256) k_ade_dm2(AdeArgs<T> m) {
  ade_dm2_cell<C, T>(m, blockIdx.x * blockDim.x + threadIdx.x);
}

// Dispersive cells of every member in one launch per class; the job's
// component selects the curl terms.
template <int CLS, class T>
__global__ void __launch_bounds__(256) k_ade_batch(const AdeJob<T>* jobs, int njobs) {
  const AdeJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const int q = (int)(blockIdx.x - J.block0) * 256 + threadIdx.x;
#define GMES_ADE_CASES(FN)                  \
  switch (J.comp) {                         \
    case 0: FN<0, T>(J.a, q); break;        \
    case 1: FN<1, T>(J.a, q); break;        \
    default: FN<2, T>(J.a, q); break;       \
  }
  if constexpr (CLS == GMES_CLS_DRUDE) { GMES_ADE_CASES(ade_drude_cell) }
  else if constexpr (CLS == GMES_CLS_DCP_ADE) { GMES_ADE_CASES(ade_dcp_cell) }
  else if constexpr (CLS == GMES_CLS_DCP_PLRC) { GMES_ADE_CASES(ade_plrc_cell) }
  else { GMES_ADE_CASES(ade_dm2_cell) }
#undef GMES_ADE_CASES
}

// ------------------------------------------------------------- running DFT
// One thread per cell of a monitor box, at the end of a step: E components
// hold their value at n = k + 1/2 and H components at n = k + 1, each
// weighted by its own row entry.  FP64 accumulation.
template <class T>
__global__ void __launch_bounds__(256) k_dft_batch(const DftJob<T>* jobs, int njobs) {
  const DftJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const int64_t q = ((int64_t)blockIdx.x - J.block0) * 256 + threadIdx.x;
  if (q >= J.ncells) return;
  const int64_t step = J.ctl->clock - 1;
  if (step % J.decim) return;
  const int64_t row = step - J.ctl->dft_step0;
  const int k = (int)(q % J.n[2]);
  const int j = (int)((q / J.n[2]) % J.n[1]);
  const int i = (int)(q / ((int64_t)J.n[2] * J.n[1]));
  const T v = gld(J.f + (kPad + ((i + J.lo[0]) * J.s1 + (j + J.lo[1])) * J.p + (k + J.lo[2]) - J.koff));
  const double vr = (double)re_of(v), vi = (double)im_of(v);
  // the job table's pointers are global memory: address them as such
  const GMES_GLOBAL double* w = (const GMES_GLOBAL double*)J.w + row * J.nfreq * 4 + (J.electric ? 0 : 2);
  // each (re, im) accumulator pair and its (re, im) compensation pair are
  // 16-byte words, read and written once per sampled step: non-temporal.
  // Compensated (Kahan) summation: c carries the low-order part that the
  // running sum lost, so the error does not grow with the number of steps.
  typedef double d2 __attribute__((ext_vector_type(2)));
  GMES_GLOBAL d2* acc = (GMES_GLOBAL d2*)J.acc + q;
  GMES_GLOBAL d2* cmp = (GMES_GLOBAL d2*)J.acc + (int64_t)J.nfreq * J.ncells + q;
#pragma unroll 4
  for (int f = 0; f < J.nfreq; ++f) {
    const double wr = w[4 * f], wi = w[4 * f + 1];
    GMES_GLOBAL d2* a = acc + f * J.ncells;
    GMES_GLOBAL d2* e = cmp + f * J.ncells;
    d2 x = __builtin_nontemporal_load(a);
    d2 c = __builtin_nontemporal_load(e);
    const double yr = (wr * vr - wi * vi) - c.x;
    const double yi = (wr * vi + wi * vr) - c.y;
    const double tr = x.x + yr, ti = x.y + yi;
    c.x = (tr - x.x) - yr;
    c.y = (ti - x.y) - yi;
    x.x = tr;
    x.y = ti;
    __builtin_nontemporal_store(x, a);
    __builtin_nontemporal_store(c, e);
  }
}

// ---------------------------------------------------------------- sources
// One thread per (component, cell) group; its ops run in table order
// (gmes/pw_source.py:74-87,106-119 point sources; :216-486 TFSF faces).
template <class T>
__device__ __forceinline__ void src_group(const SrcArgs<T>& m, int g) {
  using R = typename real_of<T>::type;
  if (g >= m.ngroups) return;
  const SrcGroup grp = m.groups[g];
  T* F = m.fa.f[grp.comp] + grp.cell;
  T v = *F;
  const int64_t row = (m.ctl->clock - 1) - m.ctl->wave_step0;
  for (int o = grp.begin; o < grp.end; ++o) {
    const DevSrcOp<R> op = m.ops[o];
    if (op.kind == GMES_SRC_SET || op.kind == GMES_SRC_SUB) {
      const double* w = m.waves + (op.wave * m.ctl->wave_n + row) * 2;
      const T wv = make_val<T>((R)w[0], (R)w[1]);
      v = op.kind == GMES_SRC_SET ? wv : vsub(v, wv);
    } else {
      const T* aux = static_cast<const T*>(m.aux[op.aux * 6 + op.aux_comp]);
      T a0 = aux[op.i0], a1 = aux[op.i1];
      if (op.wave >= 0) {  // Gaussian-beam envelope scales each sample first
        const R env = (R)m.waves[(op.wave * m.ctl->wave_n + row) * 2];
        a0 = vscl(env, a0);
        a1 = vscl(env, a1);
      }
      const T inc = vadd(vscl(op.r0, a0), vscl(op.r1, a1));
      v = op.kind == GMES_SRC_AUX_ADD ? vadd(v, vscl(op.k, inc)) : vsub(v, vscl(op.k, inc));
    }
  }
  *F = v;
}

template <class T>
__global__ void __launch_bounds__(256) k_src(SrcArgs<T> m) {
  src_group<T>(m, blockIdx.x * blockDim.x + threadIdx.x);
}

template <class T>
__global__ void __launch_bounds__(256) k_src_batch(const SrcJob<T>* jobs, int njobs) {
  const SrcJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  src_group<T>(J.a, (int)(blockIdx.x - J.block0) * 256 + threadIdx.x);
}

template <class T>
__global__ void k_probe(ProbeArgs<T> m) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m.n) return;
  const int c = m.comps[p];
  if ((c < 3) != m.electric) return;
  const T v = m.fa.f[c][m.cells[p]];
  const int64_t pos = (m.ctl->clock - 1) - m.ctl->probe_base;
  double* out = m.buf + (p * m.cap + pos) * 2;
  out[0] = (double)re_of(v);
  out[1] = (double)im_of(v);
}

// ------------------------------------------------------------- ghost planes
// Periodic / Bloch ghost copy of every active component of one field type
// (gmes/fdtd.py:811-1097).  The first copy along axis a1 and the second
// along a2 are folded into one pass: every ghost point reads only interior
// points, and the corner line gets ph2 * (ph1 * x), as the sequential
// copies produce.  Thread 0 also advances the step clock (first kernel of
// a step).
template <class T>
__device__ __forceinline__ void ghost_point(const GhostArgs<T>& m, int64_t e) {
  int t = 0;
  while (t < m.ncomp && e >= m.count[t]) { e -= m.count[t]; ++t; }
  if (t >= m.ncomp) return;
  int r = (int)(e + m.skip[t]);  // a point of one component: 32-bit index arithmetic
  const int c = m.comp[t];
  const int a1 = m.ax[t][0], a2 = m.ax[t][1], a0 = m.ax[t][2];
  const int* s = m.fa.s[c];
  const bool electric = c < 3;
  const int dst1 = electric ? s[a1] - 1 : 0, src1 = electric ? 0 : s[a1] - 1;
  const int dst2 = electric ? s[a2] - 1 : 0, src2 = electric ? 0 : s[a2] - 1;
  int ix[3];
  T v;
  const int plane1 = s[a0] * s[a2];
  T* f = m.fa.f[c];
  // consecutive points run along the free axis of higher index (the one
  // that is contiguous, or closer to contiguous, in memory)
  if (r < plane1) {
    const int fa = a0 > a2 ? a0 : a2, sa = a0 > a2 ? a2 : a0;
    ix[sa] = r / s[fa];
    ix[fa] = r - ix[sa] * s[fa];
    ix[a1] = dst1;
    int sx[3] = {ix[0], ix[1], ix[2]};
    sx[a1] = src1;
    if (ix[a2] == dst2) sx[a2] = src2;
    v = gld(f + didx(m.fa, c, sx[0], sx[1], sx[2]));
    if constexpr (is_cplx<T>::value) {
      v = cmul(m.ph[t][0], v);
      if (ix[a2] == dst2) v = cmul(m.ph[t][1], v);
    }
  } else {
    r -= plane1;
    const int w1 = s[a1] - 1;  // a1 coordinates off the first ghost plane
    int w;
    if (a1 > a0) {
      ix[a0] = r / w1;
      w = r - ix[a0] * w1;
    } else {
      w = r / s[a0];
      ix[a0] = r - w * s[a0];
    }
    ix[a1] = electric ? w : w + 1;
    ix[a2] = dst2;
    int sx[3] = {ix[0], ix[1], ix[2]};
    sx[a2] = src2;
    v = gld(f + didx(m.fa, c, sx[0], sx[1], sx[2]));
    if constexpr (is_cplx<T>::value) v = cmul(m.ph[t][1], v);
  }
  gst(f + didx(m.fa, c, ix[0], ix[1], ix[2]), v);
}

template <class T>
__global__ void __launch_bounds__(256) k_ghost(GhostArgs<T> m) {
  if (m.tick && blockIdx.x == 0 && threadIdx.x == 0) m.ctl->clock = m.ctl->clock + 1;
  ghost_point<T>(m, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Every member gets at least one block, whose thread 0 ticks its clock.
template <class T>
__global__ void __launch_bounds__(256) k_ghost_batch(const GhostJob<T>* jobs, int njobs, int tick) {
  const GhostJob<T>& J = jobs[find_job(jobs, njobs, blockIdx.x)];
  const int64_t local = (int64_t)blockIdx.x - J.block0;
  if (tick && local == 0 && threadIdx.x == 0) J.g.ctl->clock = J.g.ctl->clock + 1;
  ghost_point<T>(J.g, local * 256 + threadIdx.x);
}

// --------------------------------------------------------------- launchers
template <class T>
void launch_main2z(bool electric, const MainArgs<T>& m, int ext[3], hipStream_t st) {
  const int bj = ext[1] >= 256 ? 256 : (ext[1] >= 128 ? 128 : 64);
  dim3 block(bj), grid((ext[1] + bj - 1) / bj, (ext[0] + k2zRows - 1) / k2zRows);
  if (electric) hipLaunchKernelGGL((k_main2z<true, T>), grid, block, 0, st, m);
  else hipLaunchKernelGGL((k_main2z<false, T>), grid, block, 0, st, m);
}

template <class T>
void launch_main(bool electric, const MainArgs<T>& m, int ext[3], int bx, int by, hipStream_t st) {
  dim3 block(bx, by, 1);
  dim3 grid(m.kloop ? 1 : (ext[2] + bx - 1) / bx, (ext[1] + by - 1) / by, ext[0]);
  if (electric) hipLaunchKernelGGL((k_main<true, T>), grid, block, 0, st, m);
  else hipLaunchKernelGGL((k_main<false, T>), grid, block, 0, st, m);
}

template <class T>
void launch_main3d(bool electric, const MainArgs<T>& m, int bx, int by, int ichunk, hipStream_t st) {
  dim3 block(bx, by);
  const int n = ichunk <= 1 ? 1 : ichunk <= 2 ? 2 : ichunk <= 4 ? 4 : 8;
  dim3 grid((m.n[2] + bx - 1) / bx, (m.n[1] + by - 1) / by, (m.n[0] + n - 1) / n);
  // the UPML-prefetching variant exists for real FP32 only (the only type
  // whose kernels prefetch state words)
  // one plane per block (the default) with a record table that fits the
  // block's LDS staging (RecStage::use_lds): the LDS-only variant
  constexpr int chunks = (int)(sizeof(DevRec<typename real_of<T>::type>) / 16);
  const bool lds = m.nrec <= kRecLds && m.nrec * chunks <= bx * by;
#define GMES_L3(NN, UP, LDSO)                                                                   \
  if (n == NN) {                                                                                \
    if (electric) hipLaunchKernelGGL((k_main3d<true, T, NN, UP, LDSO>), grid, block, 0, st, m); \
    else hipLaunchKernelGGL((k_main3d<false, T, NN, UP, LDSO>), grid, block, 0, st, m);         \
    return;                                                                                     \
  }
  if constexpr (std::is_same<T, float>::value) {
    if (m.upml_mask) {
      // UPML profile rows loaded with the other operands (MainArgs::ur)
      if (m.ur && lds && n == 1) {
        if (electric) hipLaunchKernelGGL((k_main3d<true, T, 1, true, true, false, true>), grid, block, 0, st, m);
        else hipLaunchKernelGGL((k_main3d<false, T, 1, true, true, false, true>), grid, block, 0, st, m);
        return;
      }
      if (lds) GMES_L3(1, true, true)
      GMES_L3(1, true, false) GMES_L3(2, true, false) GMES_L3(4, true, false) GMES_L3(8, true, false)
    }
    // CPML profile rows loaded with the other operands (MainArgs::pp)
    if (m.pp && lds && n == 1) {
      if (electric) hipLaunchKernelGGL((k_main3d<true, T, 1, false, true, true>), grid, block, 0, st, m);
      else hipLaunchKernelGGL((k_main3d<false, T, 1, false, true, true>), grid, block, 0, st, m);
      return;
    }
  }
  if (lds) GMES_L3(1, false, true)
  GMES_L3(1, false, false) GMES_L3(2, false, false) GMES_L3(4, false, false) GMES_L3(8, false, false)
#undef GMES_L3
}

template <class T>
void launch_main3d_flat(bool electric, const MainArgs<T>& m, hipStream_t st) {
  if constexpr (std::is_same<T, float>::value) {
    const unsigned gx = ((unsigned)m.n[1] * (unsigned)m.n[2] + 255u) / 256u;
    const dim3 grid(gx * (unsigned)m.n[0]), block(256);
    constexpr int chunks = (int)(sizeof(DevRec<float>) / 16);
    const bool lds = m.nrec <= kRecLds && m.nrec * chunks <= 256;
#define GMES_LF(UP, LDSO, PP, UR)                                                                 \
  {                                                                                               \
    if (electric) hipLaunchKernelGGL((k_main3d_flat<true, T, UP, LDSO, PP, UR>), grid, block, 0, st, m); \
    else hipLaunchKernelGGL((k_main3d_flat<false, T, UP, LDSO, PP, UR>), grid, block, 0, st, m);        \
    return;                                                                                       \
  }
    if (m.upml_mask) {
      if (lds && m.ur) GMES_LF(true, true, false, true)
      if (lds) GMES_LF(true, true, false, false)
      GMES_LF(true, false, false, false)
    }
    if (m.pp && lds) GMES_LF(false, true, true, false)
    if (lds) GMES_LF(false, true, false, false)
    GMES_LF(false, false, false, false)
#undef GMES_LF
  } else {
    const unsigned gx = ((unsigned)m.n[1] * (unsigned)m.n[2] + 255u) / 256u;
    const dim3 grid(gx * (unsigned)m.n[0]), block(256);
    constexpr int chunks = (int)(sizeof(DevRec<typename real_of<T>::type>) / 16);
    const bool lds = m.nrec <= kRecLds && m.nrec * chunks <= 256;
    if (lds) {
      if (electric) hipLaunchKernelGGL((k_main3d_flat<true, T, false, true>), grid, block, 0, st, m);
      else hipLaunchKernelGGL((k_main3d_flat<false, T, false, true>), grid, block, 0, st, m);
    } else {
      if (electric) hipLaunchKernelGGL((k_main3d_flat<true, T, false, false>), grid, block, 0, st, m);
      else hipLaunchKernelGGL((k_main3d_flat<false, T, false, false>), grid, block, 0, st, m);
    }
  }
}

// Row repacking for field reads and writes of padded layouts: element k of
// row r moves from src[r * sp + k] to dst[r * dp + k].  A 2-D copy from
// the pitched device rows straight to pageable host memory ran at 0.2 GB/s
// (256^3 FP32: 0.3 s per component); packing on the device and copying
// one contiguous block is limited by the host link instead.
template <class W>
__global__ void __launch_bounds__(256) k_pitch_copy(const W* __restrict__ src, W* __restrict__ dst, int64_t total,
                                                    int n2, int sp, int dp) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  const int64_t r = t / n2;
  const int64_t k = t - r * n2;
  dst[r * dp + k] = src[r * sp + k];
}

void launch_pitch_copy(const void* src, void* dst, int64_t rows, int n2, int spitch, int dpitch, int esize,
                       hipStream_t st) {
  const int64_t total = rows * n2;
  if (total <= 0) return;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  typedef unsigned long long u64_t;
  typedef int w16_t __attribute__((ext_vector_type(4)));
  if (esize == 4)
    hipLaunchKernelGGL((k_pitch_copy<uint32_t>), grid, block, 0, st, (const uint32_t*)src, (uint32_t*)dst, total, n2,
                       spitch, dpitch);
  else if (esize == 8)
    hipLaunchKernelGGL((k_pitch_copy<u64_t>), grid, block, 0, st, (const u64_t*)src, (u64_t*)dst, total, n2, spitch,
                       dpitch);
  else
    hipLaunchKernelGGL((k_pitch_copy<w16_t>), grid, block, 0, st, (const w16_t*)src, (w16_t*)dst, total, n2, spitch,
                       dpitch);
}

template <class T>
void launch_ade(int cls, int c, const AdeArgs<T>& m, hipStream_t st) {
  dim3 block(256), grid((unsigned)((m.n + 255) / 256));
#define GMES_ADE_SWITCH(KERNEL)                                               \
  switch (c) {                                                               \
    case 0: hipLaunchKernelGGL((KERNEL<0, T>), grid, block, 0, st, m); break; \
    case 1: hipLaunchKernelGGL((KERNEL<1, T>), grid, block, 0, st, m); break; \
    default: hipLaunchKernelGGL((KERNEL<2, T>), grid, block, 0, st, m); break; \
  }
  if (cls == GMES_CLS_DRUDE) { GMES_ADE_SWITCH(k_ade_drude) }
  else if (cls == GMES_CLS_DCP_ADE) { GMES_ADE_SWITCH(k_ade_dcp) }
  else if (cls == GMES_CLS_DCP_PLRC) { GMES_ADE_SWITCH(k_ade_plrc) }
  else if (cls == GMES_CLS_DM2) { GMES_ADE_SWITCH(k_ade_dm2) }
#undef GMES_ADE_SWITCH
}

template <class T>
void launch_main_batch(bool electric, int kind, const MainJob<T>* jobs, int njobs, int64_t nblocks,
                       hipStream_t st) {
  if (njobs == 0 || nblocks == 0) return;
  const dim3 grid((unsigned)nblocks), block(256);
  if (kind == kBatch3dPP) {
    if constexpr (std::is_same<T, float>::value) {
      if (electric) hipLaunchKernelGGL((k_main3d_batch<true, T, true, true>), grid, block, 0, st, jobs, njobs);
      else hipLaunchKernelGGL((k_main3d_batch<false, T, true, true>), grid, block, 0, st, jobs, njobs);
    }
  } else if (kind == kBatch3dUR) {
    if constexpr (std::is_same<T, float>::value) {
      if (electric) hipLaunchKernelGGL((k_main3d_batch<true, T, true, false, true>), grid, block, 0, st, jobs, njobs);
      else hipLaunchKernelGGL((k_main3d_batch<false, T, true, false, true>), grid, block, 0, st, jobs, njobs);
    }
  } else if (kind == kBatch3dLds) {
    if (electric) hipLaunchKernelGGL((k_main3d_batch<true, T, true>), grid, block, 0, st, jobs, njobs);
    else hipLaunchKernelGGL((k_main3d_batch<false, T, true>), grid, block, 0, st, jobs, njobs);
  } else if (kind == kBatch3d) {
    if (electric) hipLaunchKernelGGL((k_main3d_batch<true, T, false>), grid, block, 0, st, jobs, njobs);
    else hipLaunchKernelGGL((k_main3d_batch<false, T, false>), grid, block, 0, st, jobs, njobs);
  } else if (kind == kBatch2zLds) {
    if (electric) hipLaunchKernelGGL((k_main2z_batch<true, T, true>), grid, block, 0, st, jobs, njobs);
    else hipLaunchKernelGGL((k_main2z_batch<false, T, true>), grid, block, 0, st, jobs, njobs);
  } else if (kind == kBatch2z) {
    if (electric) hipLaunchKernelGGL((k_main2z_batch<true, T, false>), grid, block, 0, st, jobs, njobs);
    else hipLaunchKernelGGL((k_main2z_batch<false, T, false>), grid, block, 0, st, jobs, njobs);
  } else {
    if (electric) hipLaunchKernelGGL((k_main_batch<true, T>), grid, block, 0, st, jobs, njobs);
    else hipLaunchKernelGGL((k_main_batch<false, T>), grid, block, 0, st, jobs, njobs);
  }
}

template <class T>
void launch_ade_batch(int cls, const AdeJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st) {
  if (njobs == 0 || nblocks == 0) return;
  const dim3 grid((unsigned)nblocks), block(256);
  if (cls == GMES_CLS_DRUDE) hipLaunchKernelGGL((k_ade_batch<GMES_CLS_DRUDE, T>), grid, block, 0, st, jobs, njobs);
  else if (cls == GMES_CLS_DCP_ADE) hipLaunchKernelGGL((k_ade_batch<GMES_CLS_DCP_ADE, T>), grid, block, 0, st, jobs, njobs);
  else if (cls == GMES_CLS_DCP_PLRC) hipLaunchKernelGGL((k_ade_batch<GMES_CLS_DCP_PLRC, T>), grid, block, 0, st, jobs, njobs);
  else if (cls == GMES_CLS_DM2) hipLaunchKernelGGL((k_ade_batch<GMES_CLS_DM2, T>), grid, block, 0, st, jobs, njobs);
}

template <class T>
void launch_dft_batch(const DftJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st) {
  if (njobs == 0 || nblocks == 0) return;
  hipLaunchKernelGGL((k_dft_batch<T>), dim3((unsigned)nblocks), dim3(256), 0, st, jobs, njobs);
}

template <class T>
void launch_src_batch(const SrcJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st) {
  if (njobs == 0 || nblocks == 0) return;
  hipLaunchKernelGGL((k_src_batch<T>), dim3((unsigned)nblocks), dim3(256), 0, st, jobs, njobs);
}

template <class T>
void launch_ghost_batch(const GhostJob<T>* jobs, int njobs, int64_t nblocks, int tick, hipStream_t st) {
  if (njobs == 0 || nblocks == 0) return;
  hipLaunchKernelGGL((k_ghost_batch<T>), dim3((unsigned)nblocks), dim3(256), 0, st, jobs, njobs, tick);
}

template <class T>
void launch_src(const SrcArgs<T>& m, hipStream_t st) {
  hipLaunchKernelGGL((k_src<T>), dim3((m.ngroups + 255) / 256), dim3(256), 0, st, m);
}

template <class T>
void launch_probe(const ProbeArgs<T>& m, hipStream_t st) {
  hipLaunchKernelGGL((k_probe<T>), dim3((m.n + 63) / 64), dim3(64), 0, st, m);
}

template <class T>
void launch_ghost(const GhostArgs<T>& m, int64_t total, hipStream_t st) {
  const int64_t blocks = total > 0 ? (total + 255) / 256 : 1;
  hipLaunchKernelGGL((k_ghost<T>), dim3((unsigned)blocks), dim3(256), 0, st, m);
}

// ---- sum of |F|^2 over a box (two passes, fixed order)
__device__ __forceinline__ double sq_of(float x) { return (double)x * (double)x; }
__device__ __forceinline__ double sq_of(double x) { return x * x; }
template <class R>
__device__ __forceinline__ double sq_of(cplx<R> x) { return sq_of(x.re) + sq_of(x.im); }

__device__ __forceinline__ double block_sum_256(double v) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

template <class T>
__global__ void __launch_bounds__(256) k_sum_sq(SqBox<T> b, double* __restrict__ partial) {
  const int base = (int)blockIdx.x * kSqBlockCells;
  double acc = 0.0;
  for (int r = 0; r < kSqBlockCells / 256; ++r) {
    const int q = base + r * 256 + (int)threadIdx.x;
    if (q < b.ncells) {
      const int k = q % b.n[2], t = q / b.n[2];
      const int j = t % b.n[1], i = t / b.n[1];
      acc = acc + sq_of(b.f[kPad + ((i + b.lo[0]) * b.s1 + (j + b.lo[1])) * b.p + (k + b.lo[2]) - b.koff]);
    }
  }
  const double tot = block_sum_256(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_sum_partials(double* __restrict__ partial, int nblocks) {
  double acc = 0.0;
  for (int q = (int)threadIdx.x; q < nblocks; q += 256) acc = acc + partial[q];
  const double tot = block_sum_256(acc);
  if (threadIdx.x == 0) partial[nblocks] = tot;
}

template <class T>
void launch_sum_sq(const SqBox<T>& b, double* partial, hipStream_t st) {
  const int nb = (int)sq_blocks(b.ncells);
  hipLaunchKernelGGL((k_sum_sq<T>), dim3(nb), dim3(256), 0, st, b, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, nb);
}

#define GMES_INSTANTIATE(T)                                                                     \
  template void launch_sum_sq<T>(const SqBox<T>&, double*, hipStream_t);                        \
  template void launch_main<T>(bool, const MainArgs<T>&, int[3], int, int, hipStream_t);       \
  template void launch_main3d<T>(bool, const MainArgs<T>&, int, int, int, hipStream_t);        \
  template void launch_main3d_flat<T>(bool, const MainArgs<T>&, hipStream_t);                 \
  template void launch_main2z<T>(bool, const MainArgs<T>&, int[3], hipStream_t);               \
  template void launch_ade<T>(int, int, const AdeArgs<T>&, hipStream_t);                       \
  template void launch_src<T>(const SrcArgs<T>&, hipStream_t);                                  \
  template void launch_probe<T>(const ProbeArgs<T>&, hipStream_t);                              \
  template void launch_ghost<T>(const GhostArgs<T>&, int64_t, hipStream_t);                     \
  template void launch_main_batch<T>(bool, int, const MainJob<T>*, int, int64_t, hipStream_t);   \
  template void launch_src_batch<T>(const SrcJob<T>*, int, int64_t, hipStream_t);               \
  template void launch_ade_batch<T>(int, const AdeJob<T>*, int, int64_t, hipStream_t);          \
  template void launch_dft_batch<T>(const DftJob<T>*, int, int64_t, hipStream_t);               \
  template void launch_ghost_batch<T>(const GhostJob<T>*, int, int64_t, int, hipStream_t);

GMES_INSTANTIATE(float)
GMES_INSTANTIATE(double)
GMES_INSTANTIATE(cplx<float>)
GMES_INSTANTIATE(cplx<double>)

}  // namespace gmes_engine
