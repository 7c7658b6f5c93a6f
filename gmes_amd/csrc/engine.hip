// C-ABI of the MI355X FDTD engine (include/gmes_engine.h).
//
// Host side of the engine: device memory for one sim, validation of every
// index a kernel will use (so a bad table becomes an error code, never an
// out-of-bounds access), grouping of source operations per cell, the
// per-step launch sequence (gmes/fdtd.py:1099-1122), HIP-graph capture of
// one step, HIP-event kernel timing and batches of sims.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <shared_mutex>
#include <string>
#include <vector>

#include "engine_kernels.h"

using namespace gmes_engine;

// Partial sums of gmes_sim_field_sum_sq_noalloc: 65536 blocks of 2048
// cells (boxes of up to 2^27 cells) and the total, 512 KiB per GPU.
constexpr int64_t kScratchPartials = 65536;
static __device__ __align__(256) double scratch[kScratchPartials + 1];

namespace {

thread_local char g_err[512];

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return set_error(e_ == hipErrorOutOfMemory ? GMES_ENOMEM : GMES_EDEVICE, "%s: %s", \
                       #expr, hipGetErrorString(e_));                                   \
  } while (0)

// Engine calls made from several threads (a served request is built on a
// worker thread while the main thread captures the batch step) must not
// run inside a stream capture: a synchronous copy, memset or free issued
// meanwhile can invalidate it.  Every entry point holds this lock shared
// (once per thread, however deeply entry points nest); a capture holds it
// exclusively, so it runs while no other thread is inside the engine.
std::shared_mutex g_capture_mu;
thread_local int tl_api_depth = 0;

struct ApiGuard {
  ApiGuard() {
    if (tl_api_depth++ == 0) g_capture_mu.lock_shared();
  }
  ~ApiGuard() {
    if (--tl_api_depth == 0) g_capture_mu.unlock_shared();
  }
};

// Exclusive for the duration of a capture; the calling thread's own shared
// hold is released first and taken back afterwards.
struct CaptureGuard {
  bool held;
  CaptureGuard() : held(tl_api_depth > 0) {
    if (held) g_capture_mu.unlock_shared();
    g_capture_mu.lock();
  }
  ~CaptureGuard() {
    g_capture_mu.unlock();
    if (held) g_capture_mu.lock_shared();
  }
};

hipError_t dev_release(void* p) { return hipFree(p); }
hipError_t device_drain() { return hipDeviceSynchronize(); }
hipError_t stream_release(hipStream_t st) { return hipStreamDestroy(st); }

const int kAx1[6] = {1, 2, 0, 1, 2, 0};
const int kAx2[6] = {2, 0, 1, 2, 0, 1};
const int kGhostOrder[6][2] = {{1, 2}, {2, 0}, {0, 1}, {1, 2}, {2, 0}, {0, 1}};

bool active(const gmes_sim* s, int c) { return (s->g.active >> c) & 1u; }
bool is3d(const gmes_sim* s) { return s->g.active == 0x3Fu; }
// 3-D sims with rows under 256 cells store their rows unpadded (gmes_sim::packed)
bool packed_rows(const gmes_grid_desc* d) { return d->active == 0x3Fu && d->nz < 256; }
// A 2-D mode with a collapsed z axis (TMz, TEz): every active component
// has rows of at most 2 elements and exactly one updated z plane (its own
// axis z has extent 1; a z extent of 2 holds one ghost plane).  Single-sim
// steps of such modes use k_main2z.
bool is_2z(const gmes_sim* s) {
  if (is3d(s)) return false;
  for (int c = 0; c < 6; ++c) {
    if (!active(s, c)) continue;
    const int64_t* sh = s->shape[c];
    if (sh[2] > 2 || (c % 3 == 2 && sh[2] != 1) || sh[0] < 2 || sh[1] < 2) return false;
  }
  return true;
}

// device element of logical (i,j,k) / of a logical C-order flat index
int64_t dev_index(const gmes_sim* s, int c, int64_t i, int64_t j, int64_t k) {
  return kPad + (i * s->shape[c][1] + j) * s->pitch[c] + k - s->koff[c];
}
int64_t dev_flat(const gmes_sim* s, int c, int64_t flat) {
  const int64_t sz = s->shape[c][2], sy = s->shape[c][1];
  return dev_index(s, c, flat / (sy * sz), (flat / sz) % sy, flat % sz);
}


int dev_alloc(gmes_sim* s, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  HIP_TRY(hipMalloc(p, bytes));
  s->bytes += (int64_t)bytes;
  return GMES_OK;
}

void dev_free(gmes_sim* s, void* p, size_t bytes) {
  if (p) {
    (void)dev_release(p);
    s->bytes -= (int64_t)bytes;
  }
}

int state_words(const gmes_mat_rec& r) {
  switch (r.cls) {
    case GMES_CLS_DRUDE: return 2 * r.n1;
    case GMES_CLS_DCP_ADE: return 1 + 2 * r.n1 + 2 * r.n2;
    case GMES_CLS_DCP_PLRC: return r.n1 + 2 * r.n2;
    case GMES_CLS_DM2: return 3 * r.n1;
    default: return 0;
  }
}

size_t real_size(const gmes_sim* s) { return s->fp64 ? 8 : 4; }

// bytes of one ADE state word of a class
size_t ade_word(const gmes_sim* s, int cls) {
  if (cls == GMES_CLS_DCP_PLRC) return 2 * real_size(s);
  if (cls == GMES_CLS_DM2) return real_size(s);
  return (size_t)s->esize;
}

void free_ade(gmes_sim* s) {
  for (int c = 0; c < 6; ++c)
    for (int k = 0; k < GMES_NCLS; ++k) {
      AdeList& L = s->ade[c][k];
      if (L.cell) (void)dev_release(L.cell);
      if (L.rec) (void)dev_release(L.rec);
      if (L.state) (void)dev_release(L.state);
      L = AdeList();
    }
}

// Byte copies of the material ids for the real-FP32 3-D main kernels,
// which then read 1 byte per cell and component instead of 2, when the
// record table has at most 256 entries (other sims and kernels read the
// 16-bit ids; for wider words the saving measured below the cost).
int upload_byte_ids(gmes_sim* s) {
  const bool narrow = is3d(s) && s->g.dtype == GMES_F32 && s->recs.size() <= 256;
  for (int c = 0; c < 6; ++c) {
    if (!narrow || !active(s, c)) {
      if (s->ids8[c]) { (void)dev_release(s->ids8[c]); s->ids8[c] = nullptr; }
      continue;
    }
    if (!s->ids8[c]) {
      int rc = dev_alloc(s, (void**)&s->ids8[c], (size_t)s->dsize[c]);
      if (rc) return rc;
    }
    const std::vector<uint16_t>& ids = s->host_ids[c];
    std::vector<uint8_t> dev((size_t)s->dsize[c], 0);
    const int64_t rows = s->shape[c][0] * s->shape[c][1], n2 = s->shape[c][2];
    for (int64_t r = 0; r < rows; ++r) {
      uint8_t* d = dev.data() + kPad - s->koff[c] + r * s->pitch[c];
      const uint16_t* h = ids.data() + r * n2;
      for (int64_t k = 0; k < n2; ++k) d[k] = (uint8_t)h[k];
    }
    HIP_TRY(hipMemcpy(s->ids8[c], dev.data(), dev.size(), hipMemcpyHostToDevice));
  }
  return GMES_OK;
}

// Coordinates along `axis` where the CPML c coefficient of profile set p is
// exactly zero for every active component: the longest such run [lo, hi),
// packed as lo | hi << 16 (0 = empty: every coordinate keeps its psi).
// There psi stays +0, and the kernels neither read nor write it.
int32_t zero_c_run(const gmes_sim* s, int p, int axis) {
  int lo = 0, hi = INT32_MAX;
  for (int c = 0; c < 6; ++c) {
    if (!active(s, c)) continue;
    const std::vector<double>& v = s->prof_host[p][c][axis];
    if (s->prof_w[p][c][axis] != 3 || v.empty()) return 0;
    const int n = (int)(v.size() / 3);
    int best_a = 0, best_b = 0;
    for (int q = 0; q < n;) {
      if (v[3 * (size_t)q + 1] != 0.0) { ++q; continue; }
      int e = q;
      while (e < n && v[3 * (size_t)e + 1] == 0.0) ++e;
      if (e - q > best_b - best_a) { best_a = q; best_b = e; }
      q = e;
    }
    lo = std::max(lo, best_a);
    hi = std::min(hi, best_b);
  }
  if (lo >= hi || hi > 0x7FFF) return 0;
  return lo | (hi << 16);
}

// Real-FP32 3-D sims whose three components of a field type all have CPML
// cells: each axis is read by two of them (E: x by Ey and Ez, y by Ex and
// Ez, z by Ex and Ey; H: x by Hy and Hz, y by Hx and Hz, z by Hx and Hy),
// at the same coordinates.  When their set-0 rows are identical, and every
// row is (b, 0, 1) inside the zero-c run of its axis, the kernels load one
// row per axis with the cell's other operands (MainArgs::pp) instead of
// two rows per component after the cell's record.
void profile_prefetch(gmes_sim* s) {
  s->pp_ok[0] = s->pp_ok[1] = false;
  if (!is3d(s) || s->g.dtype != GMES_F32 || s->nprof < 1) return;
  for (int e = 0; e < 2; ++e) {
    const int c0 = e ? 3 : 0;
    bool ok = true;
    for (int c = c0; c < c0 + 3 && ok; ++c) ok = s->pml_state[c] && s->has_cls[c][GMES_CLS_CPML];
    const int pairs[3][2] = {{c0 + 1, c0 + 2}, {c0, c0 + 2}, {c0, c0 + 1}};
    for (int a = 0; a < 3 && ok; ++a) {
      const std::vector<double>& r0 = s->prof_host[0][pairs[a][0]][a];
      const std::vector<double>& r1 = s->prof_host[0][pairs[a][1]][a];
      ok = s->prof_w[0][pairs[a][0]][a] == 3 && s->prof_w[0][pairs[a][1]][a] == 3 && !r0.empty() && r0 == r1;
      const int lo = s->prun[a] & 0xFFFF, hi = s->prun[a] >> 16;
      for (int q = lo; ok && q < hi && 3 * (size_t)q + 2 < r0.size(); ++q)
        ok = r0[3 * (size_t)q + 1] == 0.0 && r0[3 * (size_t)q + 2] == 1.0;
    }
    s->pp_ok[e] = ok;
  }
}

// Real-FP32 3-D sims: the components with UPML (and no CPML) state whose
// state word the 3-D kernels load together with the cell's other operands,
// and a box of storage coordinates (per axis lo | hi << 16) holding no UPML
// cell of those components, inside which the kernels read the array's
// guard word instead (its value is unused there).  The box is the run of
// non-UPML cells through each component's centre along each axis,
// intersected over the components, and is kept only if no cell inside it
// is UPML; otherwise the kernels load the word after the cell's record.
void upml_box(gmes_sim* s) {
  s->upml_pre = 0;
  for (int a = 0; a < 3; ++a) s->ubox[a] = 0;
  if (!is3d(s) || s->g.dtype != GMES_F32) return;
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0x7FFF, 0x7FFF, 0x7FFF};
  unsigned mask = 0;
  for (int c = 0; c < 6; ++c) {
    if (!active(s, c) || !s->pml_state[c] || s->has_cls[c][GMES_CLS_CPML] || !s->has_cls[c][GMES_CLS_UPML]) continue;
    mask |= 1u << c;
    const int64_t* n = s->shape[c];
    const std::vector<uint16_t>& ids = s->host_ids[c];
    for (int a = 0; a < 3; ++a) {
      int64_t ix[3] = {n[0] / 2, n[1] / 2, n[2] / 2};
      auto upml_at = [&](int64_t v) {
        ix[a] = v;
        return s->recs[ids[(ix[0] * n[1] + ix[1]) * n[2] + ix[2]]].cls == GMES_CLS_UPML;
      };
      const int64_t mid = n[a] / 2;
      if (upml_at(mid)) return;
      int64_t l = mid, h = mid + 1;
      while (l > 0 && !upml_at(l - 1)) --l;
      while (h < n[a] && !upml_at(h)) ++h;
      lo[a] = std::max(lo[a], l);
      hi[a] = std::min(hi[a], h);
    }
  }
  if (!mask) return;
  for (int a = 0; a < 3; ++a)
    if (lo[a] >= hi[a]) return;
  for (int c = 0; c < 6; ++c) {
    if (!((mask >> c) & 1u)) continue;
    const int64_t* n = s->shape[c];
    const std::vector<uint16_t>& ids = s->host_ids[c];
    for (int64_t i = lo[0]; i < std::min(hi[0], n[0]); ++i)
      for (int64_t j = lo[1]; j < std::min(hi[1], n[1]); ++j)
        for (int64_t k = lo[2]; k < std::min(hi[2], n[2]); ++k)
          if (s->recs[ids[(i * n[1] + j) * n[2] + k]].cls == GMES_CLS_UPML) return;
  }
  s->upml_pre = mask;
  for (int a = 0; a < 3; ++a) s->ubox[a] = (int32_t)(lo[a] | (hi[a] << 16));
}

// Real-FP32 3-D sims whose three components of a field type all get their
// UPML state prefetched (upml_pre): each axis is a curl axis of two of
// them (E: x of Ey and Ez, y of Ex and Ez, z of Ex and Ey; H: x of Hy and
// Hz, y of Hx and Hz, z of Hx and Hy), which take c1..c4 at the same
// coordinates, and the own axis of the third, which takes c5, c6.  When
// the two set-0 rows of each axis are identical, the kernels load that
// row and the owner's row with the cell's other operands (MainArgs::ur)
// instead of three rows per component after the cell's record.
void upml_rows(gmes_sim* s) {
  s->ur_ok[0] = s->ur_ok[1] = false;
  if (!is3d(s) || s->g.dtype != GMES_F32 || s->nprof < 1) return;
  for (int e = 0; e < 2; ++e) {
    const int c0 = e ? 3 : 0;
    bool ok = true;
    for (int c = c0; c < c0 + 3 && ok; ++c) ok = (s->upml_pre >> c) & 1u;
    const int pairs[3][2] = {{c0 + 1, c0 + 2}, {c0, c0 + 2}, {c0, c0 + 1}};
    for (int a = 0; a < 3 && ok; ++a) {
      const std::vector<double>& r0 = s->prof_host[0][pairs[a][0]][a];
      const std::vector<double>& r1 = s->prof_host[0][pairs[a][1]][a];
      ok = s->prof_w[0][pairs[a][0]][a] == 6 && s->prof_w[0][pairs[a][1]][a] == 6 &&
           s->prof_w[0][c0 + a][a] == 6 && !r0.empty() && r0 == r1 && !s->prof_host[0][c0 + a][a].empty();
    }
    s->ur_ok[e] = ok;
  }
}

template <class R>
int upload_tables(gmes_sim* s) {
  std::vector<DevRec<R>> dr(s->recs.size());
  for (size_t i = 0; i < s->recs.size(); ++i) {
    const gmes_mat_rec& r = s->recs[i];
    dr[i].cls = r.cls;
    dr[i].n1 = r.n1;
    dr[i].n2 = r.n2;
    dr[i].prof = r.prof;
    dr[i].coef_off = r.coef_off;
    if (r.cls == GMES_CLS_CPML && r.prof >= 0 && r.prof < s->nprof) {
      dr[i].n1 = zero_c_run(s, r.prof, 0);
      dr[i].n2 = zero_c_run(s, r.prof, 1);
      dr[i].coef_off = zero_c_run(s, r.prof, 2);
    }
    dr[i].c0 = (R)(s->g.dt / r.eps_mu);  // dt / eps_inf in FP64, rounded once
    dr[i].eps_mu = (R)r.eps_mu;
    dr[i].inv_eps = (R)(1.0 / r.eps_mu);
    dr[i].val_re = (R)r.value[0];
    dr[i].val_im = (R)r.value[1];
  }
  std::vector<R> cf(s->coef.size());
  for (size_t i = 0; i < cf.size(); ++i) cf[i] = (R)s->coef[i];
  // profiles into one buffer with offsets
  std::vector<int64_t> off(64 * 18, 0);
  std::vector<R> pr;
  for (int p = 0; p < s->nprof; ++p)
    for (int c = 0; c < 6; ++c)
      for (int a = 0; a < 3; ++a) {
        off[(p * 6 + c) * 3 + a] = (int64_t)pr.size();
        // FP32 tables hold 1/kappa in the CPML kappa column (the FP32 kernels
        // multiply); FP64 tables hold the reference's values unchanged.
        const std::vector<double>& src = s->prof_host[p][c][a];
        const bool recip = !s->fp64 && s->prof_w[p][c][a] == 3;
        for (size_t q = 0; q < src.size(); ++q)
          pr.push_back((R)(recip && q % 3 == 2 ? 1.0 / src[q] : src[q]));
      }
  if (pr.size() > (size_t)INT32_MAX) return set_error(GMES_EINVAL, "profile pool too large");
  for (int c = 0; c < 6; ++c)
    for (int a = 0; a < 3; ++a) s->prof_off0[c][a] = off[c * 3 + a];
  if (pr.empty()) pr.push_back(R(0));
  if (cf.empty()) cf.push_back(R(0));
  int rc;
  if ((rc = dev_alloc(s, &s->d_recs, dr.size() * sizeof(DevRec<R>)))) return rc;
  if ((rc = dev_alloc(s, &s->d_coef, cf.size() * sizeof(R)))) return rc;
  if ((rc = dev_alloc(s, &s->d_prof, pr.size() * sizeof(R)))) return rc;
  if ((rc = dev_alloc(s, (void**)&s->d_prof_off, off.size() * sizeof(int64_t)))) return rc;
  HIP_TRY(hipMemcpy(s->d_recs, dr.data(), dr.size() * sizeof(DevRec<R>), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_coef, cf.data(), cf.size() * sizeof(R), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_prof, pr.data(), pr.size() * sizeof(R), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_prof_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  return GMES_OK;
}

template <class R>
int upload_sources(gmes_sim* s) {
  // stable grouping by (phase, comp, cell); E groups first
  std::vector<int64_t> order(s->ops.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    const gmes_src_op &a = s->ops[x], &b = s->ops[y];
    const int pa = a.comp >= 3, pb = b.comp >= 3;
    if (pa != pb) return pa < pb;
    if (a.comp != b.comp) return a.comp < b.comp;
    return a.cell < b.cell;
  });
  std::vector<DevSrcOp<R>> dops(order.size());
  std::vector<SrcGroup> groups;
  s->ngroups_e = s->ngroups_h = 0;
  for (size_t n = 0; n < order.size(); ++n) {
    const gmes_src_op& o = s->ops[order[n]];
    DevSrcOp<R>& d = dops[n];
    d.kind = o.kind; d.wave = o.wave; d.aux = o.aux; d.aux_comp = o.aux_comp;
    d.i0 = o.i0; d.i1 = o.i1; d.k = (R)o.k; d.r0 = (R)o.r0; d.r1 = (R)o.r1;
    if (o.kind == GMES_SRC_AUX_ADD || o.kind == GMES_SRC_AUX_SUB) {
      d.i0 = dev_flat(s->aux[o.aux], o.aux_comp, o.i0);
      d.i1 = dev_flat(s->aux[o.aux], o.aux_comp, o.i1);
    }
    const int64_t dcell = dev_flat(s, o.comp, o.cell);
    if (groups.empty() || groups.back().comp != o.comp || groups.back().cell != dcell) {
      SrcGroup g;
      g.comp = o.comp; g.cell = dcell; g.begin = (int32_t)n; g.end = (int32_t)n + 1; g.pad_ = 0;
      groups.push_back(g);
      if (o.comp < 3) s->ngroups_e++; else s->ngroups_h++;
    } else {
      groups.back().end = (int32_t)n + 1;
    }
  }
  if (s->d_ops) { dev_free(s, s->d_ops, 0); s->d_ops = nullptr; }
  if (s->d_groups) { dev_free(s, s->d_groups, 0); s->d_groups = nullptr; }
  int rc;
  if ((rc = dev_alloc(s, &s->d_ops, std::max<size_t>(1, dops.size()) * sizeof(DevSrcOp<R>)))) return rc;
  if ((rc = dev_alloc(s, (void**)&s->d_groups, std::max<size_t>(1, groups.size()) * sizeof(SrcGroup)))) return rc;
  if (!dops.empty())
    HIP_TRY(hipMemcpy(s->d_ops, dops.data(), dops.size() * sizeof(DevSrcOp<R>), hipMemcpyHostToDevice));
  if (!groups.empty())
    HIP_TRY(hipMemcpy(s->d_groups, groups.data(), groups.size() * sizeof(SrcGroup), hipMemcpyHostToDevice));
  return GMES_OK;
}

// ------------------------------------------------------------ launching
template <class T>
FieldArgs<T> field_args(const gmes_sim* s) {
  FieldArgs<T> a;
  for (int c = 0; c < 6; ++c) {
    a.f[c] = static_cast<T*>(s->f[c]);
    for (int d = 0; d < 3; ++d) a.s[c][d] = (int)s->shape[c][d];
    a.p[c] = s->pitch[c];
    a.koff[c] = s->koff[c];
  }
  a.active = s->g.active;
  return a;
}

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

void time_begin(gmes_sim* s, int fam, hipStream_t st) {
  if (!s->timing.on) return;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, st);
  s->timing.pending.push_back({fam, {a, b}});
}

void time_end(gmes_sim* s, hipStream_t st) {
  if (!s->timing.on) return;
  (void)hipEventRecord(s->timing.pending.back().second.second, st);
}

// Ghost-copy arguments of one field type; total = ghost points.
template <class T>
GhostArgs<T> ghost_args(const gmes_sim* s, bool electric, int tick, int64_t& total) {
  using R = typename real_of<T>::type;
  GhostArgs<T> g{};
  g.fa = field_args<T>(s);
  g.ncomp = 0;
  total = 0;
  for (int c = electric ? 0 : 3; c < (electric ? 3 : 6); ++c) {
    if (!active(s, c)) continue;
    const int t = g.ncomp++;
    const int a1 = kGhostOrder[c][0], a2 = kGhostOrder[c][1], a0 = c % 3;
    g.comp[t] = c;
    g.ax[t][0] = a1; g.ax[t][1] = a2; g.ax[t][2] = a0;
    const int64_t p1 = s->shape[c][a0] * s->shape[c][a2], p2 = s->shape[c][a0] * (s->shape[c][a1] - 1);
    g.count[t] = p1 + p2;
    g.skip[t] = 0;
    if (!electric && s->fuse_hz && s->shape[c][2] == 2) {  // the z plane is the E kernel's
      if (a2 == 2) g.count[t] = p1;
      else if (a1 == 2) { g.skip[t] = p1; g.count[t] = p2; }
    }
    g.ph[t][0] = cplx<R>{(R)s->g.phase[c][a1][0], (R)s->g.phase[c][a1][1]};
    g.ph[t][1] = cplx<R>{(R)s->g.phase[c][a2][0], (R)s->g.phase[c][a2][1]};
    total += g.count[t];
  }
  g.tick = tick;
  g.ctl = reinterpret_cast<Ctl*>(s->d_clock);
  return g;
}

template <class T>
void enqueue_ghost(gmes_sim* s, bool electric, int tick, hipStream_t st) {
  int64_t total = 0;
  GhostArgs<T> g = ghost_args<T>(s, electric, tick, total);
  time_begin(s, 4, st);
  launch_ghost<T>(g, total, st);
  time_end(s, st);
}

// Main-kernel arguments of one field type; ext = the largest storage extent
// of its components.  False when the field type has no active component.
template <class T>
bool main_args(const gmes_sim* s, bool electric, MainArgs<T>& m, int ext[3]) {
  using R = typename real_of<T>::type;
  m = MainArgs<T>{};
  m.fa = field_args<T>(s);
  ext[0] = ext[1] = ext[2] = 0;
  bool any = false;
  for (int c = electric ? 0 : 3; c < (electric ? 3 : 6); ++c) {
    if (!active(s, c)) continue;
    any = true;
    for (int d = 0; d < 3; ++d) ext[d] = std::max<int>(ext[d], (int)s->shape[c][d]);
  }
  if (!any) return false;
  m.kloop = !is3d(s) && ext[2] <= 2;
  m.kext = ext[2];
  m.cpml_mask = 0;
  for (int c = 0; c < 6; ++c)
    if (s->pml_state[c] && s->has_cls[c][GMES_CLS_CPML]) m.cpml_mask |= 1u << c;
  for (int a = 0; a < 3; ++a) m.prun[a] = s->prun[a];
  m.upml_mask = s->upml_pre;
  for (int a = 0; a < 3; ++a) m.ubox[a] = s->ubox[a];
  m.hzg = electric && s->fuse_hz;
  if (m.hzg) {
    m.hzph[0] = cplx<R>{(R)s->g.phase[3][2][0], (R)s->g.phase[3][2][1]};
    m.hzph[1] = cplx<R>{(R)s->g.phase[4][2][0], (R)s->g.phase[4][2][1]};
    m.hyxph = cplx<R>{(R)s->g.phase[4][0][0], (R)s->g.phase[4][0][1]};
  }
  for (int c = 0; c < 6; ++c)
    for (int a = 0; a < 3; ++a) m.po0[c][a] = (int)s->prof_off0[c][a];
  m.ur = 0;
  if (s->ur_ok[electric ? 0 : 1]) {
    // shared rows: x of the 2nd component, y and z of the 1st; own rows:
    // axis a of component a
    const int c0 = electric ? 0 : 3;
    m.ur = 1;
    m.urs[0] = m.po0[c0 + 1][0];
    m.urs[1] = m.po0[c0][1];
    m.urs[2] = m.po0[c0][2];
    for (int a = 0; a < 3; ++a) m.uro[a] = m.po0[c0 + a][a];
  }
  m.pp = 0;
  if (s->pp_ok[electric ? 0 : 1]) {
    // axis x: 2nd component (Ey / Hy), axis y and z: 1st component (Ex / Hx)
    const int c0 = electric ? 0 : 3;
    m.pp = 1;
    m.pprow[0] = m.po0[c0 + 1][0];
    m.pprow[1] = m.po0[c0][1];
    m.pprow[2] = m.po0[c0][2];
  }
  m.narrow = 1;
  for (int c = 0; c < 6; ++c) {
    m.ids[c] = s->ids[c];
    m.ids8[c] = s->ids8[c];
    if (active(s, c) && !s->ids8[c]) m.narrow = 0;
    m.pml[c] = s->pml_state[c];
    m.dsize[c] = (int)s->dsize[c];
  }
  m.recs = static_cast<const DevRec<R>*>(s->d_recs);
  m.prof = static_cast<const R*>(s->d_prof);
  m.prof_off = s->d_prof_off;
  m.d[0] = (R)s->g.dx; m.d[1] = (R)s->g.dy; m.d[2] = (R)s->g.dz;
  m.inv_d[0] = (R)(1.0 / s->g.dx); m.inv_d[1] = (R)(1.0 / s->g.dy); m.inv_d[2] = (R)(1.0 / s->g.dz);
  m.nrec = (int)s->recs.size();
  m.n[0] = s->g.nx; m.n[1] = s->g.ny; m.n[2] = s->g.nz;
  m.fuse = electric && s->fuse_e;
  for (int c = 0; c < 3; ++c)
    for (int t = 0; t < 2; ++t) {
      const int a = kGhostOrder[c][t];
      m.gax[c][t] = a;
      m.gph[c][t] = cplx<R>{(R)s->g.phase[c][a][0], (R)s->g.phase[c][a][1]};
    }
  return true;
}


// Tile (bx along k, by along j) of the main kernels: the sim's own setting
// (gmes_sim_set_tile, gmes_sim_autotune) or the default.  3-D: (j,k) tiles
// of one i plane, 128 x 2 for long rows; on packed rows (gmes_sim::packed)
// 256 x 1 means flat blocks of 256 consecutive (j, k) points of a plane,
// the default there.  Other modes: a k x j tile of one i plane.
void tile_of(const gmes_sim* s, const int ext[3], int& bx, int& by) {
  bx = s->tile_bx;
  by = s->tile_by;
  if (bx > 0 && by > 0 && (is3d(s) || (bx <= next_pow2(ext[2]) && ext[2] > 2))) return;
  if (is3d(s) && s->packed) {
    bx = 256;
    by = 1;
  } else if (is3d(s)) {
    // real FP32: whole 256-cell rows per block when they divide the row
    // (256^3: E launch 0.1708 to 0.1675 ms, H unchanged; FP64 0.5% slower
    // that way), else 128 x 2
    bx = s->esize == 4 && s->g.nz % 256 == 0 ? 256 : std::min(128, next_pow2(s->g.nz));
    by = std::max(1, 256 / bx);
  } else if (ext[2] <= 2) {
    bx = 1;  // short rows: a thread per (i, j) walks k (MainArgs::kloop)
    by = std::min(256, next_pow2(ext[1]));
  } else {
    bx = std::min(64, next_pow2(ext[2]));
    by = std::max(1, std::min(256 / bx, next_pow2(ext[1])));
  }
}

template <class T>
void launch_main_tile(const gmes_sim* s, bool electric, const MainArgs<T>& m, int ext[3], int bx, int by,
                      hipStream_t st) {
  if (is3d(s) && s->packed && bx == 256 && by == 1) {
    launch_main3d_flat<T>(electric, m, st);
  } else if (is3d(s)) {
    launch_main3d<T>(electric, m, bx, by, s->tile_ic > 0 ? s->tile_ic : 1, st);
  } else {
    launch_main<T>(electric, m, ext, bx, by, st);
  }
}

template <class T>
void enqueue_main(gmes_sim* s, bool electric, hipStream_t st) {
  MainArgs<T> m;
  int ext[3];
  if (!main_args<T>(s, electric, m, ext)) return;
  time_begin(s, electric ? 0 : 1, st);
  if (is_2z(s)) {
    launch_main2z<T>(electric, m, ext, st);
  } else {
    int bx, by;
    tile_of(s, ext, bx, by);
    launch_main_tile<T>(s, electric, m, ext, bx, by, st);
  }
  time_end(s, st);
}

template <class T>
AdeArgs<T> ade_args(const gmes_sim* s, int c, int cls) {
  using R = typename real_of<T>::type;
  const AdeList& L = s->ade[c][cls];
  AdeArgs<T> a{};
  a.fa = field_args<T>(s);
  a.cell = L.cell; a.rec = L.rec; a.state = L.state; a.n = (int)L.n;
  a.recs = static_cast<const DevRec<R>*>(s->d_recs);
  a.coef = static_cast<const R*>(s->d_coef);
  a.d[0] = (R)s->g.dx; a.d[1] = (R)s->g.dy; a.d[2] = (R)s->g.dz;
  a.dt = (R)s->g.dt;
  a.ctl = reinterpret_cast<Ctl*>(s->d_clock);
  return a;
}

template <class T>
void enqueue_ade(gmes_sim* s, hipStream_t st) {
  for (int c = 0; c < 3; ++c)
    for (int cls = GMES_CLS_DRUDE; cls < GMES_NCLS; ++cls) {
      if (s->ade[c][cls].n == 0) continue;
      const AdeArgs<T> a = ade_args<T>(s, c, cls);
      time_begin(s, 2, st);
      launch_ade<T>(cls, c, a, st);
      time_end(s, st);
    }
}

template <class T>
SrcArgs<T> src_args(const gmes_sim* s, bool electric) {
  using R = typename real_of<T>::type;
  SrcArgs<T> a{};
  a.fa = field_args<T>(s);
  a.groups = s->d_groups + (electric ? 0 : s->ngroups_e);
  a.ngroups = electric ? s->ngroups_e : s->ngroups_h;
  a.ops = static_cast<const DevSrcOp<R>*>(s->d_ops);
  a.waves = s->d_waves;
  a.aux = static_cast<void* const*>(s->d_aux_ptrs);
  a.ctl = reinterpret_cast<Ctl*>(s->d_clock);
  return a;
}

template <class T>
void enqueue_src(gmes_sim* s, bool electric, hipStream_t st) {
  const SrcArgs<T> a = src_args<T>(s, electric);
  if (a.ngroups == 0) return;
  time_begin(s, 3, st);
  launch_src<T>(a, st);
  time_end(s, st);
}

template <class T>
void enqueue_probe(gmes_sim* s, bool electric, hipStream_t st) {
  if (s->probes.empty()) return;
  ProbeArgs<T> a{};
  a.fa = field_args<T>(s);
  a.cells = s->d_probe_cells;
  a.comps = s->d_probe_comps;
  a.n = (int)s->probes.size();
  a.electric = electric;
  a.buf = s->d_probe_buf;
  a.cap = s->probe_cap;
  a.ctl = reinterpret_cast<Ctl*>(s->d_clock);
  launch_probe<T>(a, st);
}

// DFT monitor jobs of a sim, first block numbered block0.
template <class T>
void dft_jobs(const gmes_sim* s, std::vector<DftJob<T>>& out, int64_t& block0) {
  for (const auto& d : s->dfts) {
    DftJob<T> J{};
    J.f = static_cast<T*>(s->f[d.comp]);
    J.s1 = (int)s->shape[d.comp][1];
    J.p = s->pitch[d.comp];
    J.koff = s->koff[d.comp];
    J.comp = d.comp;
    for (int a = 0; a < 3; ++a) { J.lo[a] = (int)d.lo[a]; J.n[a] = (int)d.n[a]; }
    J.nfreq = d.nfreq;
    J.decim = d.decim;
    J.electric = d.comp < 3;
    J.ncells = d.ncells;
    J.acc = d.acc;
    J.w = s->d_dft_w;
    J.ctl = reinterpret_cast<Ctl*>(s->d_clock);
    J.block0 = block0;
    block0 += (d.ncells + 255) / 256;
    out.push_back(J);
  }
}

template <class T>
int upload_dft_jobs_t(gmes_sim* s) {
  std::vector<DftJob<T>> jobs;
  int64_t nb = 0;
  dft_jobs<T>(s, jobs, nb);
  s->dft_blocks = nb;
  if (jobs.empty()) return GMES_OK;
  // capacity for 16 monitors: adding monitors rarely reallocates (a free
  // would wait for every kernel on the device)
  const int64_t need = std::max<int64_t>(16, (int64_t)jobs.size());
  if (need > s->dft_jobs_cap) {
    if (s->d_dft_jobs) HIP_TRY(dev_release(s->d_dft_jobs));
    s->d_dft_jobs = nullptr;
    HIP_TRY(hipMalloc(&s->d_dft_jobs, (size_t)need * sizeof(DftJob<T>)));
    s->dft_jobs_cap = need;
  }
  HIP_TRY(hipMemcpy(s->d_dft_jobs, jobs.data(), jobs.size() * sizeof(DftJob<T>), hipMemcpyHostToDevice));
  return GMES_OK;
}

int upload_dft_jobs(gmes_sim* s) {
  switch (s->g.dtype) {
    case GMES_F32: return upload_dft_jobs_t<float>(s);
    case GMES_F64: return upload_dft_jobs_t<double>(s);
    case GMES_C64: return upload_dft_jobs_t<cplx<float>>(s);
    default: return upload_dft_jobs_t<cplx<double>>(s);
  }
}

template <class T>
void enqueue_dft(gmes_sim* s, hipStream_t st) {
  if (s->dfts.empty()) return;
  launch_dft_batch<T>(static_cast<const DftJob<T>*>(s->d_dft_jobs), (int)s->dfts.size(), s->dft_blocks, st);
}

void enqueue_step(gmes_sim* s, hipStream_t st);

// One whole step of each auxiliary (TFSF) grid of s, between the E and H
// half steps of s.
void enqueue_aux_steps(gmes_sim* s, hipStream_t st) {
  for (int i = 0; i < s->naux; ++i) {
    if (!s->aux[i]) continue;
    time_begin(s, 5, st);
    enqueue_step(s->aux[i], st);
    time_end(s, st);
  }
}

template <class T>
void enqueue_step_t(gmes_sim* s, hipStream_t st) {
  enqueue_ghost<T>(s, false, 1, st);
  enqueue_main<T>(s, true, st);
  enqueue_ade<T>(s, st);
  enqueue_src<T>(s, true, st);
  enqueue_probe<T>(s, true, st);
  enqueue_aux_steps(s, st);
  if (!s->fuse_e) enqueue_ghost<T>(s, true, 0, st);
  enqueue_main<T>(s, false, st);
  enqueue_src<T>(s, false, st);
  enqueue_probe<T>(s, false, st);
  enqueue_dft<T>(s, st);
}

void enqueue_step(gmes_sim* s, hipStream_t st) {
  switch (s->g.dtype) {
    case GMES_F32: enqueue_step_t<float>(s, st); break;
    case GMES_F64: enqueue_step_t<double>(s, st); break;
    case GMES_C64: enqueue_step_t<cplx<float>>(s, st); break;
    default: enqueue_step_t<cplx<double>>(s, st); break;
  }
}

void advance_host_clock(gmes_sim* s, int64_t n) {
  s->step += n;
  s->probe_n = s->step - s->probe_base;
  for (int i = 0; i < s->naux; ++i)
    if (s->aux[i]) advance_host_clock(s->aux[i], n);
}

// Waits for all work that may touch s: its own stream, for an aux sim the
// stream of the sim that advances it, and the stream of s's batch.
hipError_t sync_exec(gmes_sim* s) {
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess && s->parent) e = sync_exec(s->parent);
  if (e == hipSuccess && s->batch_stream) e = hipStreamSynchronize(s->batch_stream);
  return e;
}

// Checks that every step in [s->step, s->step + n) is covered by the
// waveform window and the probe buffer (recursively for aux sims).
int check_windows(gmes_sim* s, int64_t n) {
  bool waves = false;
  for (const auto& o : s->ops)
    if (o.kind == GMES_SRC_SET || o.kind == GMES_SRC_SUB || o.wave >= 0) waves = true;
  if (waves && (s->step < s->wave_step0 || s->step + n > s->wave_step0 + s->wave_n))
    return set_error(GMES_EINVAL, "waveforms cover steps [%lld, %lld), need [%lld, %lld)",
                     (long long)s->wave_step0, (long long)(s->wave_step0 + s->wave_n),
                     (long long)s->step, (long long)(s->step + n));
  if (!s->dfts.empty() && (s->step < s->dft_step0 || s->step + n > s->dft_step0 + s->dft_n))
    return set_error(GMES_EINVAL, "DFT weights cover steps [%lld, %lld), need [%lld, %lld)",
                     (long long)s->dft_step0, (long long)(s->dft_step0 + s->dft_n), (long long)s->step,
                     (long long)(s->step + n));
  for (int i = 0; i < s->naux; ++i) {
    if (!s->aux[i]) continue;
    int rc = check_windows(s->aux[i], n);
    if (rc) return rc;
  }
  return GMES_OK;
}

int ensure_probe_capacity(gmes_sim* s, int64_t n) {
  for (int i = 0; i < s->naux; ++i) {
    if (!s->aux[i]) continue;
    int rc = ensure_probe_capacity(s->aux[i], n);
    if (rc) return rc;
  }
  if (s->probes.empty()) return GMES_OK;
  const int64_t need = (s->step - s->probe_base) + n;
  if (need <= s->probe_cap) return GMES_OK;
  int64_t cap = std::max<int64_t>(1024, s->probe_cap);
  while (cap < need) cap *= 2;
  double* nb = nullptr;
  int rc = dev_alloc(s, (void**)&nb, (size_t)cap * s->probes.size() * 2 * sizeof(double));
  if (rc) return rc;
  HIP_TRY(sync_exec(s));
  if (s->d_probe_buf) {
    const int64_t used = s->step - s->probe_base;
    for (size_t p = 0; p < s->probes.size(); ++p)
      HIP_TRY(hipMemcpy(nb + p * cap * 2, s->d_probe_buf + p * s->probe_cap * 2, used * 2 * sizeof(double),
                        hipMemcpyDeviceToDevice));
    dev_free(s, s->d_probe_buf, (size_t)s->probe_cap * s->probes.size() * 2 * sizeof(double));
  }
  s->d_probe_buf = nb;
  s->probe_cap = cap;
  s->graph_dirty = true;
  ++s->gen;
  if (s->parent) {
    s->parent->graph_dirty = true;
    ++s->parent->gen;
  }
  return GMES_OK;
}

void destroy_graph(gmes_sim* s) {
  if (s->graph_exec) {
    (void)hipGraphExecDestroy(s->graph_exec);
    s->graph_exec = nullptr;
  }
  s->graph_dirty = true;
  ++s->gen;
  // the parent's graph holds this aux sim's kernels and their arguments
  if (s->parent) destroy_graph(s->parent);
}

int build_graph(gmes_sim* s) {
  destroy_graph(s);
  hipGraph_t graph;
  {
    CaptureGuard capture;
    HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(s, s->stream);
    HIP_TRY(hipStreamEndCapture(s->stream, &graph));
  }
  hipError_t e = hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return set_error(GMES_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
  s->graph_dirty = false;
  return GMES_OK;
}

int collect_timing(gmes_sim* s) {
  for (auto& p : s->timing.pending) {
    float ms = 0.f;
    HIP_TRY(hipEventSynchronize(p.second.second));
    HIP_TRY(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    s->timing.ms[p.first] += ms;
    s->timing.launches[p.first] += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  s->timing.pending.clear();
  return GMES_OK;
}

bool any_graph_blocker(gmes_sim* s) { return s->timing.on; }

int do_step(gmes_sim* s, int64_t n) {
  if (!s->committed) return set_error(GMES_EINVAL, "sim is not committed");
  if (n < 0) return set_error(GMES_EINVAL, "negative step count");
  if (s->parent) return set_error(GMES_EINVAL, "an auxiliary grid is advanced by its parent");
  if (s->aux_lost) return set_error(GMES_EINVAL, "an auxiliary grid of this sim was destroyed");
  if (s->batch_stream) HIP_TRY(hipStreamSynchronize(s->batch_stream));
  int rc = check_windows(s, n);
  if (rc) return rc;
  if ((rc = ensure_probe_capacity(s, n))) return rc;
  HIP_TRY(hipSetDevice(s->device));
  if (s->use_graph && !any_graph_blocker(s)) {
    if (s->graph_dirty || !s->graph_exec)
      if ((rc = build_graph(s))) return rc;
    for (int64_t t = 0; t < n; ++t) HIP_TRY(hipGraphLaunch(s->graph_exec, s->stream));
  } else {
    for (int64_t t = 0; t < n; ++t) enqueue_step(s, s->stream);
    HIP_TRY(hipGetLastError());
  }
  advance_host_clock(s, n);
  s->probe_n = s->step - s->probe_base;
  if (s->timing.on && s->timing.pending.size() > 4096) return collect_timing(s);
  return GMES_OK;
}

// ---- tile autotuner
// Times the E main kernel of a sim under each candidate tile and keeps the
// fastest, per size class (device, storage type, mode, log2 of the
// extents) for the life of the process.  The fields and CPML state the
// trial launches touch are saved first and restored afterwards, so tuning
// leaves the simulation unchanged.
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tune_cache;

int ceil_log2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

template <class T>
int autotune_t(gmes_sim* s, int& best_bx, int& best_by) {
  MainArgs<T> m;
  int ext[3];
  best_bx = best_by = 0;
  if (!main_args<T>(s, true, m, ext)) return GMES_OK;
  if (!is3d(s)) return GMES_OK;  // the 2-D/1-D tile follows each field type's rows
  char key[96];
  snprintf(key, sizeof key, "%d/%d/%x/%d/%d/%d", s->device, s->g.dtype, s->g.active, ceil_log2(ext[0]),
           ceil_log2(ext[1]), ceil_log2(ext[2]));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      best_bx = it->second.first;
      best_by = it->second.second;
      return GMES_OK;
    }
  }
  std::vector<int> cand;
  const int row = next_pow2(ext[2]);
  for (int bx = 1; bx <= 256; bx *= 2) {
    if (bx == 256 && s->packed) { cand.push_back(bx); continue; }  // flat blocks
    if (bx > std::max(row, 1) && bx > 1) continue;          // wider than a row: idle lanes
    if (row >= 32 && bx < 32) continue;                      // narrower than a 128-byte line
    cand.push_back(bx);
  }
  if (cand.empty()) cand.push_back(std::min(256, row));
  HIP_TRY(sync_exec(s));
  // save what the E kernel writes: E fields and E-side PML state
  std::vector<std::pair<void*, std::pair<void*, size_t>>> saved;
  for (int c = 0; c < 3; ++c) {
    if (!active(s, c)) continue;
    const size_t fb = (size_t)s->dsize[c] * s->esize;
    void* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, fb));
    HIP_TRY(hipMemcpy(tmp, s->f[c], fb, hipMemcpyDeviceToDevice));
    saved.push_back({tmp, {s->f[c], fb}});
    if (s->pml_state[c]) {
      const int pw = s->has_cls[c][GMES_CLS_CPML] ? 2 : 1;
      const size_t pb = (size_t)pw * s->dsize[c] * s->esize;
      HIP_TRY(hipMalloc(&tmp, pb));
      HIP_TRY(hipMemcpy(tmp, s->pml_state[c], pb, hipMemcpyDeviceToDevice));
      saved.push_back({tmp, {s->pml_state[c], pb}});
    }
  }
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  float best = 1e30f;
  const int reps = 5;
  for (int bx : cand) {
    const int by = std::max(1, 256 / bx);
    launch_main_tile<T>(s, true, m, ext, bx, by, s->stream);
    HIP_TRY(hipEventRecord(e0, s->stream));
    for (int r = 0; r < reps; ++r) launch_main_tile<T>(s, true, m, ext, bx, by, s->stream);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_bx = bx;
      best_by = by;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto& x : saved) {
    HIP_TRY(hipMemcpy(x.second.first, x.first, x.second.second, hipMemcpyDeviceToDevice));
    HIP_TRY(dev_release(x.first));
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = {best_bx, best_by};
  return GMES_OK;
}

// Whether the 3-D (or 2-D collapsed-z) E kernel may write the E ghost
// planes itself instead of the separate copy that runs later in the step
// (after the dispersive updates, the E sources and the E probes).  The results are the same when
// none of those touches a plane involved in the copy: no dispersive cell
// on a source plane (plane 0 of a transverse axis), no E source operation
// on a source or ghost plane, no E probe on a ghost plane (it would read
// the ghost before the copy).  gmes_sim_set_ghost_fusion(sim, 0) turns it off.
bool on_ghost_copy_plane(const gmes_sim* s, int c, int64_t x, bool source, bool ghost) {
  const int64_t sy = s->shape[c][1], sz = s->shape[c][2];
  const int64_t ix[3] = {x / (sy * sz), (x / sz) % sy, x % sz};
  for (int t = 0; t < 2; ++t) {
    const int a = kGhostOrder[c][t];
    if ((source && ix[a] == 0) || (ghost && ix[a] == s->shape[c][a] - 1)) return true;
  }
  return false;
}

//
// In a 2-D mode with a collapsed z axis where Ez is the only E component
// (TMz), the z ghost plane (k = 0) of Hx and Hy is read by no E update
// (Ez reads k = 1), and nothing runs between the H ghost copy and the E
// kernel; the 2-D E kernel then writes that plane itself from k = 1, with
// the copy's phases and corner order, and the H copy leaves it out.
void decide_fusion(gmes_sim* s) {
  const bool prev = s->fuse_e, prev_hz = s->fuse_hz;
  bool f = s->allow_fuse && (is3d(s) || is_2z(s)) && s->committed && !s->ade_on_ghost_source;
  for (size_t n = 0; f && n < s->ops.size(); ++n)
    if (s->ops[n].comp < 3 && on_ghost_copy_plane(s, s->ops[n].comp, s->ops[n].cell, true, true)) f = false;
  for (size_t n = 0; f && n < s->probes.size(); ++n)
    if (s->probes[n].comp < 3 && on_ghost_copy_plane(s, s->probes[n].comp, s->probes[n].cell, false, true)) f = false;
  s->fuse_e = f;
  bool hz = s->allow_fuse && s->committed && is_2z(s) && active(s, 2) && !active(s, 0) && !active(s, 1);
  bool any = false;
  for (int c = 3; hz && c < 5; ++c) {
    if (!active(s, c) || s->shape[c][2] != 2) continue;
    any = true;
    // the E kernel's threads span Ez's (i, j) extent
    if (s->shape[c][0] > s->shape[2][0] || s->shape[c][1] > s->shape[2][1]) hz = false;
  }
  s->fuse_hz = hz && any;
  if (f != prev || s->fuse_hz != prev_hz) destroy_graph(s);
}

int ctl_write(gmes_sim* s) {
  Ctl c;
  HIP_TRY(sync_exec(s));
  HIP_TRY(hipMemcpyAsync(&c, s->d_clock, sizeof c, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(sync_exec(s));
  c.wave_step0 = s->wave_step0;
  c.wave_n = s->wave_n;
  c.probe_base = s->probe_base;
  c.dft_step0 = s->dft_step0;
  c.dft_n = s->dft_n;
  HIP_TRY(hipMemcpyAsync(s->d_clock, &c, sizeof c, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(sync_exec(s));
  return GMES_OK;
}

// ---- sum of |F|^2 over a box
// No-allocation variant: the partial sums live in this file-scope device
// array (one copy per GPU, held until the process ends).  Host code never
// passes the array's name to a kernel or copy: it uses the device address
// that hipGetSymbolAddress returns, looked up once per device with that
// device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];   // serialises no-allocation calls per device
double* g_scratch_ptr[kMaxDevices] = {};

static int sq_check(const gmes_sim* s, int32_t c, const int64_t lo[3], const int64_t hi[3], const double* out,
                    int64_t& ncells) {
  if (!s || !lo || !hi || !out) return set_error(GMES_EINVAL, "null argument");
  if (c < 0 || c > 5) return set_error(GMES_EINVAL, "bad component");
  ncells = 1;
  for (int a = 0; a < 3; ++a) {
    if (lo[a] < 0 || hi[a] > s->shape[c][a] || hi[a] <= lo[a]) return set_error(GMES_ERANGE, "box out of bounds");
    ncells *= hi[a] - lo[a];
  }
  return GMES_OK;
}

template <class T>
static void sq_launch(const gmes_sim* s, int c, const int64_t lo[3], const int64_t hi[3], double* partial) {
  SqBox<T> b;
  b.f = static_cast<const T*>(s->f[c]);
  b.s1 = (int)s->shape[c][1];
  b.p = s->pitch[c];
  b.koff = s->koff[c];
  b.ncells = 1;
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = (int)lo[a];
    b.n[a] = (int)(hi[a] - lo[a]);
    b.ncells *= b.n[a];
  }
  launch_sum_sq<T>(b, partial, s->stream);
}

// Runs the two passes on the sim's stream into partial[0 .. nblocks] and
// copies partial[nblocks] to *out.
static int sq_run(gmes_sim* s, int32_t c, const int64_t lo[3], const int64_t hi[3], double* out, double* partial,
                  int64_t ncells) {
  HIP_TRY(sync_exec(s));  // steps of a batch or a parent write these fields on their streams
  switch (s->g.dtype) {
    case GMES_F32: sq_launch<float>(s, c, lo, hi, partial); break;
    case GMES_F64: sq_launch<double>(s, c, lo, hi, partial); break;
    case GMES_C64: sq_launch<cplx<float>>(s, c, lo, hi, partial); break;
    default: sq_launch<cplx<double>>(s, c, lo, hi, partial); break;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, partial + sq_blocks(ncells), sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return GMES_OK;
}

}  // namespace

// =============================================================== C-ABI
extern "C" {

const char* gmes_last_error(void) { return g_err; }
int gmes_abi_version(void) { return GMES_ABI_VERSION; }

int gmes_device_count(int32_t* count) {
  ApiGuard api;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return GMES_OK;
}

int gmes_sim_create(const gmes_grid_desc* desc, gmes_sim** out) {
  ApiGuard api;
  if (!desc || !out) return set_error(GMES_EINVAL, "null argument");
  if (desc->dtype < GMES_F32 || desc->dtype > GMES_C128) return set_error(GMES_ETYPE, "bad dtype");
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return set_error(GMES_EINVAL, "grid sizes must be positive");
  // Kernels index a component with 32-bit offsets (and ADE cell lists hold
  // 32-bit cells): refuse any component of 2^31 device elements or more,
  // before touching the device.
  for (int c = 0; c < 6; ++c) {
    if (!((desc->active >> c) & 1u)) continue;
    const int64_t ext[3] = {(int64_t)desc->nx + (c % 3 != 0), (int64_t)desc->ny + (c % 3 != 1),
                            (int64_t)desc->nz + (c % 3 != 2)};
    const int64_t row = ext[2] >= 32 && !packed_rows(desc) ? (ext[2] + 31) / 32 * 32 : ext[2];
    const double elems = 2.0 * kPad + (double)ext[0] * (double)ext[1] * (double)row;
    if (elems >= 2147483648.0)
      return set_error(GMES_EINVAL, "component %d would hold %.0f device elements; the limit is 2^31 - 1", c, elems);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(GMES_EDEVICE, "no GPU visible");
  if (desc->device < 0 || desc->device >= ndev) return set_error(GMES_EINVAL, "bad device ordinal");
  HIP_TRY(hipSetDevice(desc->device));
  gmes_sim* s = new gmes_sim();
  s->g = *desc;
  s->device = desc->device;
  s->complex_ = desc->dtype == GMES_C64 || desc->dtype == GMES_C128;
  s->fp64 = desc->dtype == GMES_F64 || desc->dtype == GMES_C128;
  s->esize = (s->fp64 ? 8 : 4) * (s->complex_ ? 2 : 1);
  const int64_t n[3] = {desc->nx, desc->ny, desc->nz};
  // 3-D sims with rows under 256 cells store their rows back to back:
  // with 128-byte lines, padding short rows to 32 words left a part-used
  // line at every row end (the batched 64-instance workload's 3-D main
  // launches: 1.112 to 1.049 ms (E), 1.058 to 1.002 ms (H) unpadded; the
  // c64 dispersive cell's E launch 0.558 to 0.491 ms); their single-sim
  // launches then take flat blocks (tile_of).
  s->packed = packed_rows(desc);
  for (int c = 0; c < 6; ++c) {
    if (!active(s, c)) {
      s->shape[c][0] = s->shape[c][1] = s->shape[c][2] = 1;
    } else {
      for (int a = 0; a < 3; ++a) s->shape[c][a] = n[a] + 1;
      s->shape[c][c % 3] = n[c % 3];
    }
    s->ncell[c] = s->shape[c][0] * s->shape[c][1] * s->shape[c][2];
    const int64_t s2 = s->shape[c][2];
    const bool padded = s2 >= 32 && !s->packed;
    s->pitch[c] = (int32_t)(padded ? (s2 + 31) / 32 * 32 : s2);
    s->koff[c] = padded && (c == GMES_HX || c == GMES_HY) ? 1 : 0;
    s->dsize[c] = 2 * kPad + s->shape[c][0] * s->shape[c][1] * s->pitch[c];
    if (s->dsize[c] >= (int64_t(1) << 31)) {
      delete s;
      return set_error(GMES_EINVAL, "component %d has more than 2^31 cells", c);
    }
  }
  int rc = GMES_OK;
  for (int c = 0; c < 6 && !rc; ++c) {
    rc = dev_alloc(s, &s->f[c], (size_t)(s->dsize[c] * s->esize));
    if (!rc && hipMemset(s->f[c], 0, (size_t)(s->dsize[c] * s->esize)) != hipSuccess)
      rc = set_error(GMES_EDEVICE, "hipMemset failed");
  }
  if (!rc) rc = dev_alloc(s, (void**)&s->d_clock, sizeof(Ctl));
  if (!rc && hipMemset(s->d_clock, 0, sizeof(Ctl)) != hipSuccess) rc = set_error(GMES_EDEVICE, "hipMemset failed");
  if (!rc) rc = dev_alloc(s, &s->d_aux_ptrs, 48 * sizeof(void*));
  if (!rc && hipStreamCreate(&s->stream) != hipSuccess) rc = set_error(GMES_EDEVICE, "hipStreamCreate failed");
  if (rc) {
    gmes_sim_destroy(s);
    return rc;
  }
  *out = s;
  return GMES_OK;
}

int gmes_sim_destroy(gmes_sim* s) {
  ApiGuard api;
  if (!s) return GMES_OK;
  // a batch's job tables and graph hold this sim's buffers
  if (s->batch_stream) return set_error(GMES_EINVAL, "sim is a batch member: destroy the batch first");
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  // an aux sim is advanced on its parent's stream: drain that too.  The
  // parent's source operations read this grid, so the parent refuses to
  // step from now on.
  if (gmes_sim* p = s->parent) {
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->batch_stream) (void)hipStreamSynchronize(p->batch_stream);
    for (int i = 0; i < p->naux; ++i)
      if (p->aux[i] == s) p->aux[i] = nullptr;
    p->aux_lost = true;
    ++p->gen;
    destroy_graph(p);
  }
  for (int i = 0; i < s->naux; ++i)
    if (s->aux[i]) s->aux[i]->parent = nullptr;
  destroy_graph(s);
  collect_timing(s);
  for (int c = 0; c < 6; ++c) {
    if (s->f[c]) (void)dev_release(s->f[c]);
    if (s->ids[c]) (void)dev_release(s->ids[c]);
    if (s->ids8[c]) (void)dev_release(s->ids8[c]);
    if (s->pml_state[c]) (void)dev_release(s->pml_state[c]);
  }
  free_ade(s);
  for (auto& d : s->dfts)
    if (d.acc) (void)dev_release(d.acc);
  void* bufs[] = {s->d_recs, s->d_coef, s->d_prof, s->d_prof_off, s->d_ops, s->d_groups, s->d_waves,
                  s->d_aux_ptrs, s->d_probe_cells, s->d_probe_comps, s->d_probe_buf, s->d_clock,
                  s->d_dft_w, s->d_dft_jobs, s->d_sq, s->d_stage};
  for (void* b : bufs)
    if (b) (void)dev_release(b);
  if (s->stream) (void)stream_release(s->stream);
  delete s;
  return GMES_OK;
}

int gmes_sim_shape(const gmes_sim* s, int32_t c, int64_t shape[3]) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5) return set_error(GMES_EINVAL, "bad component");
  for (int a = 0; a < 3; ++a) shape[a] = s->shape[c][a];
  return GMES_OK;
}

int gmes_sim_footprint(const gmes_sim* s, int64_t* bytes) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  *bytes = s->bytes;
  return GMES_OK;
}

int gmes_sim_set_materials(gmes_sim* s, const gmes_mat_rec* recs, int32_t nrec, const double* coef,
                           int64_t ncoef) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (nrec < 1 || nrec > 65536) return set_error(GMES_EINVAL, "record count must be in [1, 65536]");
  if (ncoef < 0) return set_error(GMES_EINVAL, "negative coefficient count");
  s->recs.assign(recs, recs + nrec);
  s->coef.assign(coef, coef + ncoef);
  for (int r = 0; r < nrec; ++r) {
    const gmes_mat_rec& m = recs[r];
    if (m.cls < 0 || m.cls >= GMES_NCLS) return set_error(GMES_EINVAL, "record %d: bad class", r);
    if (m.n1 < 0 || m.n2 < 0) return set_error(GMES_EINVAL, "record %d: bad pole count", r);
    int64_t need = 0;
    switch (m.cls) {
      case GMES_CLS_DRUDE: need = 3 * (int64_t)m.n1 + 3; break;
      case GMES_CLS_DCP_ADE: need = 3 * (int64_t)m.n1 + 5 * (int64_t)m.n2 + 4; break;
      case GMES_CLS_DCP_PLRC: need = 3 * (int64_t)m.n1 + 6 * (int64_t)m.n2 + 3; break;
      case GMES_CLS_DM2: need = 2 * (int64_t)m.n1 + 6; break;
      default: break;
    }
    if (need && (m.coef_off < 0 || m.coef_off + need > ncoef))
      return set_error(GMES_EINVAL, "record %d: coefficients out of the pool", r);
    if (m.cls == GMES_CLS_DM2 && s->complex_) return set_error(GMES_EINVAL, "Dm2 class supports real fields only");
    if (m.cls == GMES_CLS_DM2 && (m.n1 < 1 || m.n1 > kDm2Max))
      return set_error(GMES_EINVAL, "record %d: DM2 supports 1..%d transitions", r, kDm2Max);
    if ((m.cls == GMES_CLS_CPML || m.cls == GMES_CLS_UPML) && (m.prof < 0 || m.prof >= 64))
      return set_error(GMES_EINVAL, "record %d: bad profile set", r);
    if (m.cls != GMES_CLS_DUMMY && m.cls != GMES_CLS_CONST && !(m.eps_mu != 0.0))
      return set_error(GMES_EINVAL, "record %d: zero permittivity/permeability", r);
  }
  s->committed = false;
  return GMES_OK;
}

int gmes_sim_set_material_ids(gmes_sim* s, int32_t c, const uint16_t* ids, int64_t count) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5) return set_error(GMES_EINVAL, "bad component");
  if (count != s->ncell[c]) return set_error(GMES_EINVAL, "id count %lld != cells %lld", (long long)count,
                                             (long long)s->ncell[c]);
  s->host_ids[c].assign(ids, ids + count);
  if (!s->ids[c]) {
    int rc = dev_alloc(s, (void**)&s->ids[c], (size_t)s->dsize[c] * sizeof(uint16_t));
    if (rc) return rc;
  }
  // the device layout (pads included) is assembled on the host and copied
  // in one contiguous transfer: a strided 2-D copy would wait for every
  // kernel running on the device, including other sims' batches
  std::vector<uint16_t> dev((size_t)s->dsize[c], 0);
  const int64_t rows = s->shape[c][0] * s->shape[c][1], n2 = s->shape[c][2];
  for (int64_t r = 0; r < rows; ++r)
    memcpy(dev.data() + kPad - s->koff[c] + r * s->pitch[c], ids + r * n2, (size_t)n2 * sizeof(uint16_t));
  HIP_TRY(hipMemcpy(s->ids[c], dev.data(), dev.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  s->committed = false;
  return GMES_OK;
}

int gmes_sim_set_profile(gmes_sim* s, int32_t p, int32_t c, int32_t axis, const double* vals, int32_t n,
                         int32_t width) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (p < 0 || p >= 64 || c < 0 || c > 5 || axis < 0 || axis > 2)
    return set_error(GMES_EINVAL, "bad profile address");
  if (n != s->shape[c][axis] || (width != 3 && width != 6)) return set_error(GMES_EINVAL, "bad profile size");
  s->prof_host[p][c][axis].assign(vals, vals + (size_t)n * width);
  s->prof_w[p][c][axis] = width;
  s->nprof = std::max(s->nprof, p + 1);
  s->committed = false;
  return GMES_OK;
}

int gmes_sim_commit(gmes_sim* s) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (s->recs.empty()) return set_error(GMES_EINVAL, "materials not set");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  free_ade(s);
  s->ade_on_ghost_source = false;
  for (int c = 0; c < 6; ++c) {
    for (int k = 0; k < GMES_NCLS; ++k) s->has_cls[c][k] = false;
    if (!active(s, c)) continue;
    const std::vector<uint16_t>& ids = s->host_ids[c];
    if ((int64_t)ids.size() != s->ncell[c]) return set_error(GMES_EINVAL, "ids of component %d not set", c);
    const int64_t s0 = s->shape[c][0], sy = s->shape[c][1], sz = s->shape[c][2];
    std::vector<std::vector<int32_t>> cells(GMES_NCLS);
    std::vector<std::vector<uint16_t>> recs(GMES_NCLS);
    std::vector<int> words(GMES_NCLS, 0);
    // records in use
    std::vector<int64_t> count(s->recs.size(), 0);
    for (int64_t x = 0; x < s->ncell[c]; ++x) {
      const uint16_t id = ids[x];
      if (id >= s->recs.size()) return set_error(GMES_ERANGE, "material id out of range");
      ++count[id];
    }
    bool any_ade = false;
    for (size_t id = 0; id < s->recs.size(); ++id) {
      if (!count[id]) continue;
      const gmes_mat_rec& r = s->recs[id];
      s->has_cls[c][r.cls] = true;
      if (r.cls >= GMES_CLS_DRUDE) {
        if (c >= 3) return set_error(GMES_EINVAL, "dispersive class on a magnetic component");
        any_ade = true;
        words[r.cls] = std::max(words[r.cls], state_words(r));
      }
      if (r.cls == GMES_CLS_CPML || r.cls == GMES_CLS_UPML) {
        const int w = r.cls == GMES_CLS_CPML ? 3 : 6;
        for (int a = 0; a < 3; ++a)
          if ((int64_t)s->prof_host[r.prof][c][a].size() != s->shape[c][a] * w)
            return set_error(GMES_EINVAL, "profile %d of component %d axis %d missing", r.prof, c, a);
      }
    }
    // Ghost planes must be Dummy: a kernel never reads past them
    // (gmes/fdtd.py:365-366 and the analogous lines).  E: the last plane
    // of each transverse axis; H: plane 0.
    const int64_t n3[3] = {s0, sy, sz};
    for (int a = 0; a < 3; ++a) {
      if (a == c % 3) continue;
      const int64_t g = c < 3 ? n3[a] - 1 : 0;
      int64_t ix[3];
      for (ix[0] = 0; ix[0] < s0; ++ix[0])
        for (ix[1] = 0; ix[1] < sy; ++ix[1]) {
          if (a == 0 && ix[0] != g) break;
          if (a == 1 && ix[1] != g) continue;
          for (ix[2] = (a == 2 ? g : 0); ix[2] < (a == 2 ? g + 1 : sz); ++ix[2]) {
            const int64_t x = (ix[0] * sy + ix[1]) * sz + ix[2];
            if (s->recs[ids[x]].cls != GMES_CLS_DUMMY)
              return set_error(GMES_EINVAL, "component %d: ghost cell %lld is not Dummy", c, (long long)x);
          }
        }
    }
    if (any_ade)
      for (int64_t x = 0; x < s->ncell[c]; ++x) {
        const gmes_mat_rec& r = s->recs[ids[x]];
        if (r.cls >= GMES_CLS_DRUDE) {
          cells[r.cls].push_back((int32_t)x);
          recs[r.cls].push_back(ids[x]);
          if (c < 3 && on_ghost_copy_plane(s, c, x, true, false)) s->ade_on_ghost_source = true;
        }
      }
    // full-grid CPML (2 words) / UPML (1 word) state
    const int pw = s->has_cls[c][GMES_CLS_CPML] ? 2 : (s->has_cls[c][GMES_CLS_UPML] ? 1 : 0);
    if (s->pml_state[c]) { dev_free(s, s->pml_state[c], 0); s->pml_state[c] = nullptr; }
    if (pw) {
      const size_t b = (size_t)pw * s->dsize[c] * s->esize;
      int rc = dev_alloc(s, &s->pml_state[c], b);
      if (rc) return rc;
      HIP_TRY(hipMemset(s->pml_state[c], 0, b));
    }
    for (int k = GMES_CLS_DRUDE; k < GMES_NCLS; ++k) {
      AdeList& L = s->ade[c][k];
      L.n = (int64_t)cells[k].size();
      if (!L.n) continue;
      L.nstate = words[k];
      HIP_TRY(hipMalloc((void**)&L.cell, L.n * sizeof(int32_t)));
      HIP_TRY(hipMalloc((void**)&L.rec, L.n * sizeof(uint16_t)));
      const size_t sb = std::max<size_t>(1, (size_t)L.nstate * L.n * ade_word(s, k));
      HIP_TRY(hipMalloc(&L.state, sb));
      s->bytes += L.n * 6 + (int64_t)sb;
      HIP_TRY(hipMemcpy(L.cell, cells[k].data(), L.n * sizeof(int32_t), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(L.rec, recs[k].data(), L.n * sizeof(uint16_t), hipMemcpyHostToDevice));
      HIP_TRY(hipMemset(L.state, 0, sb));
    }
  }
  if (s->d_recs) { (void)dev_release(s->d_recs); s->d_recs = nullptr; }
  if (s->d_coef) { (void)dev_release(s->d_coef); s->d_coef = nullptr; }
  if (s->d_prof) { (void)dev_release(s->d_prof); s->d_prof = nullptr; }
  if (s->d_prof_off) { (void)dev_release(s->d_prof_off); s->d_prof_off = nullptr; }
  int rc = s->fp64 ? upload_tables<double>(s) : upload_tables<float>(s);
  if (rc) return rc;
  if ((rc = upload_byte_ids(s))) return rc;
  // the coordinate runs where no CPML record's psi can be nonzero
  for (int a = 0; a < 3; ++a) {
    int lo = 0, hi = 0x7FFF;
    bool any = false;
    for (const gmes_mat_rec& r : s->recs) {
      if (r.cls != GMES_CLS_CPML) continue;
      const int32_t run = (r.prof >= 0 && r.prof < s->nprof) ? zero_c_run(s, r.prof, a) : 0;
      any = true;
      lo = std::max(lo, (int)(run & 0xFFFF));
      hi = std::min(hi, (int)(run >> 16));
    }
    s->prun[a] = (any && lo < hi) ? (lo | (hi << 16)) : 0;
  }
  profile_prefetch(s);
  upml_box(s);
  upml_rows(s);
  s->committed = true;
  destroy_graph(s);
  decide_fusion(s);
  return GMES_OK;
}

int gmes_sim_set_sources(gmes_sim* s, const gmes_src_op* ops, int64_t nops, int32_t nwaves) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (nops < 0 || nwaves < 0) return set_error(GMES_EINVAL, "negative count");
  if (nops > (1 << 30)) return set_error(GMES_EINVAL, "too many source operations");
  for (int64_t i = 0; i < nops; ++i) {
    const gmes_src_op& o = ops[i];
    if (o.comp < 0 || o.comp > 5 || !active(s, o.comp)) return set_error(GMES_EINVAL, "source on an inactive component");
    if (o.cell < 0 || o.cell >= s->ncell[o.comp]) return set_error(GMES_ERANGE, "field index is out of bounds");
    if (o.kind == GMES_SRC_SET || o.kind == GMES_SRC_SUB) {
      if (o.wave < 0 || o.wave >= nwaves) return set_error(GMES_EINVAL, "bad waveform row");
    } else if (o.kind == GMES_SRC_AUX_ADD || o.kind == GMES_SRC_AUX_SUB) {
      if (o.aux < 0 || o.aux >= s->naux) return set_error(GMES_EINVAL, "auxiliary grid %d not attached", o.aux);
      if (o.wave < -1 || o.wave >= nwaves) return set_error(GMES_EINVAL, "bad envelope row");
      if (o.aux_comp < 0 || o.aux_comp > 5) return set_error(GMES_EINVAL, "bad aux component");
      const gmes_sim* a = s->aux[o.aux];
      if (!((a->g.active >> o.aux_comp) & 1u)) return set_error(GMES_EINVAL, "inactive aux component");
      if (o.i0 < 0 || o.i1 < 0 || o.i0 >= a->ncell[o.aux_comp] || o.i1 >= a->ncell[o.aux_comp])
        return set_error(GMES_ERANGE, "aux sample index out of bounds");
    } else {
      return set_error(GMES_EINVAL, "bad source kind");
    }
  }
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  s->ops.assign(ops, ops + nops);
  s->nwaves = nwaves;
  int rc = s->fp64 ? upload_sources<double>(s) : upload_sources<float>(s);
  destroy_graph(s);
  decide_fusion(s);
  return rc;
}

int gmes_sim_set_waveforms(gmes_sim* s, int64_t step0, int64_t nsteps, const double* values) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (nsteps < 0) return set_error(GMES_EINVAL, "negative step count");
  HIP_TRY(hipSetDevice(s->device));
  const int64_t need = std::max<int64_t>(1, (int64_t)s->nwaves * nsteps * 2);
  HIP_TRY(sync_exec(s));
  if (need > s->wave_cap) {
    if (s->d_waves) dev_free(s, s->d_waves, (size_t)s->wave_cap * sizeof(double));
    s->d_waves = nullptr;
    int rc = dev_alloc(s, (void**)&s->d_waves, (size_t)need * sizeof(double));
    if (rc) return rc;
    s->wave_cap = need;
    destroy_graph(s);
  }
  if (s->nwaves && nsteps)
    HIP_TRY(hipMemcpy(s->d_waves, values, (size_t)s->nwaves * nsteps * 2 * sizeof(double), hipMemcpyHostToDevice));
  s->wave_step0 = step0;
  s->wave_n = nsteps;
  return ctl_write(s);
}

int gmes_sim_attach_aux(gmes_sim* s, gmes_sim* aux, int32_t* slot) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (!aux || aux == s) return set_error(GMES_EINVAL, "bad auxiliary sim");
  if (s->naux >= 8) return set_error(GMES_EINVAL, "too many auxiliary grids");
  if (aux->g.dtype != s->g.dtype) return set_error(GMES_ETYPE, "aux dtype mismatch");
  if (aux->device != s->device) return set_error(GMES_EINVAL, "aux sim on another device");
  if (aux->naux) return set_error(GMES_EINVAL, "an auxiliary grid cannot have its own");
  if (aux->parent) return set_error(GMES_EINVAL, "auxiliary grid already attached");
  s->aux[s->naux] = aux;
  aux->parent = s;
  void* ptrs[6];
  for (int c = 0; c < 6; ++c) ptrs[c] = aux->f[c];
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemcpy(static_cast<void**>(s->d_aux_ptrs) + s->naux * 6, ptrs, sizeof ptrs, hipMemcpyHostToDevice));
  *slot = s->naux++;
  destroy_graph(s);
  return GMES_OK;
}

int gmes_sim_add_probe(gmes_sim* s, int32_t c, int64_t cell, int32_t* id) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5 || !active(s, c) || cell < 0 || cell >= s->ncell[c])
    return set_error(GMES_ERANGE, "probe index out of bounds");
  if (s->step != s->probe_base) return set_error(GMES_EINVAL, "read pending probe values first");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  s->probes.push_back({c, cell});
  std::vector<int64_t> cells;
  std::vector<int32_t> comps;
  for (auto& p : s->probes) { cells.push_back(dev_flat(s, p.comp, p.cell)); comps.push_back(p.comp); }
  if (s->d_probe_cells) (void)dev_release(s->d_probe_cells);
  if (s->d_probe_comps) (void)dev_release(s->d_probe_comps);
  if (s->d_probe_buf) (void)dev_release(s->d_probe_buf);
  s->d_probe_buf = nullptr;
  s->probe_cap = 0;
  HIP_TRY(hipMalloc((void**)&s->d_probe_cells, cells.size() * sizeof(int64_t)));
  HIP_TRY(hipMalloc((void**)&s->d_probe_comps, comps.size() * sizeof(int32_t)));
  HIP_TRY(hipMemcpy(s->d_probe_cells, cells.data(), cells.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_probe_comps, comps.data(), comps.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  *id = (int32_t)s->probes.size() - 1;
  destroy_graph(s);
  decide_fusion(s);
  return GMES_OK;
}

int gmes_sim_read_probe(gmes_sim* s, int32_t id, double* out, int64_t max, int64_t* n) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (id < 0 || id >= (int32_t)s->probes.size()) return set_error(GMES_EINVAL, "bad probe id");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  std::vector<int64_t>& cur = s->probe_read;
  cur.resize(s->probes.size(), 0);
  const int64_t avail = s->probe_n - cur[id];
  const int64_t m = std::min(avail, max);
  if (m > 0)
    HIP_TRY(hipMemcpy(out, s->d_probe_buf + (id * s->probe_cap + cur[id]) * 2, m * 2 * sizeof(double),
                      hipMemcpyDeviceToHost));
  cur[id] += m;
  *n = m;
  bool all = true;
  for (size_t p = 0; p < cur.size(); ++p) all = all && cur[p] == s->probe_n;
  if (all) {
    s->probe_base = s->step;
    s->probe_n = 0;
    std::fill(cur.begin(), cur.end(), 0);
    return ctl_write(s);
  }
  return GMES_OK;
}

int gmes_sim_add_dft(gmes_sim* s, int32_t c, const int64_t lo[3], const int64_t hi[3], int32_t nfreq,
                     int32_t decim, int32_t* id) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5 || !active(s, c)) return set_error(GMES_EINVAL, "DFT of an inactive component");
  if (nfreq < 1 || decim < 1) return set_error(GMES_EINVAL, "nfreq and decim must be positive");
  if (!s->dfts.empty() && nfreq != s->dft_nfreq)
    return set_error(GMES_EINVAL, "all DFT monitors of a sim share one frequency list");
  gmes_sim::Dft d{};
  d.comp = c;
  d.ncells = 1;
  for (int a = 0; a < 3; ++a) {
    if (lo[a] < 0 || hi[a] > s->shape[c][a] || hi[a] <= lo[a]) return set_error(GMES_ERANGE, "DFT box out of bounds");
    d.lo[a] = lo[a];
    d.n[a] = hi[a] - lo[a];
    d.ncells *= d.n[a];
  }
  d.nfreq = nfreq;
  d.decim = decim;
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  // sums [nfreq][ncells][re, im], then their compensations, same layout
  const size_t bytes = 2 * (size_t)nfreq * d.ncells * 2 * sizeof(double);
  HIP_TRY(hipMalloc((void**)&d.acc, bytes));
  HIP_TRY(hipMemset(d.acc, 0, bytes));
  s->dfts.push_back(d);
  s->dft_nfreq = nfreq;
  s->bytes += (int64_t)bytes;
  *id = (int32_t)s->dfts.size() - 1;
  destroy_graph(s);
  return upload_dft_jobs(s);
}

int gmes_sim_set_dft_weights(gmes_sim* s, int64_t step0, int64_t nsteps, const double* w) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (s->dfts.empty()) return set_error(GMES_EINVAL, "no DFT monitor");
  if (nsteps < 0) return set_error(GMES_EINVAL, "negative step count");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  const int64_t need = std::max<int64_t>(1, nsteps * s->dft_nfreq * 4);
  if (need > s->dft_w_cap) {
    if (s->d_dft_w) HIP_TRY(dev_release(s->d_dft_w));
    s->d_dft_w = nullptr;
    HIP_TRY(hipMalloc((void**)&s->d_dft_w, (size_t)need * sizeof(double)));
    s->dft_w_cap = need;
    int rc = upload_dft_jobs(s);  // the jobs hold the weight pointer
    if (rc) return rc;
    ++s->gen;
  }
  if (nsteps) HIP_TRY(hipMemcpy(s->d_dft_w, w, (size_t)nsteps * s->dft_nfreq * 4 * sizeof(double), hipMemcpyHostToDevice));
  s->dft_step0 = step0;
  s->dft_n = nsteps;
  return ctl_write(s);
}

int gmes_sim_read_dft(gmes_sim* s, int32_t id, double* out, int64_t count) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (id < 0 || id >= (int32_t)s->dfts.size()) return set_error(GMES_EINVAL, "bad DFT id");
  const auto& d = s->dfts[id];
  if (count != (int64_t)d.nfreq * d.ncells * 2) return set_error(GMES_EINVAL, "size mismatch");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  HIP_TRY(hipMemcpy(out, d.acc, (size_t)count * sizeof(double), hipMemcpyDeviceToHost));
  return GMES_OK;
}

// The staging buffer of field reads and writes of padded rows, at least
// `bytes` long (grown on demand; the sim's stream is idle when it grows).
static int stage_reserve(gmes_sim* s, int64_t bytes) {
  if (bytes <= s->stage_cap) return GMES_OK;
  HIP_TRY(sync_exec(s));
  if (s->d_stage) dev_free(s, s->d_stage, (size_t)s->stage_cap);
  s->d_stage = nullptr;
  s->stage_cap = 0;
  int rc = dev_alloc(s, &s->d_stage, (size_t)bytes);
  if (rc) return rc;
  s->stage_cap = bytes;
  return GMES_OK;
}

int gmes_sim_write_field(gmes_sim* s, int32_t c, const void* host, int64_t bytes) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5) return set_error(GMES_EINVAL, "bad component");
  if (bytes != s->ncell[c] * s->esize) return set_error(GMES_EINVAL, "size mismatch");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));  // steps queued on the sim's (or its batch's) stream come first
  if (s->pitch[c] == s->shape[c][2] && s->koff[c] == 0) {
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(s->f[c]) + (size_t)kPad * s->esize, host, (size_t)bytes,
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(sync_exec(s));
    return GMES_OK;
  }
  // one contiguous upload, then rows moved to their pitched places
  int rc = stage_reserve(s, bytes);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(s->d_stage, host, (size_t)bytes, hipMemcpyHostToDevice, s->stream));
  launch_pitch_copy(s->d_stage, static_cast<char*>(s->f[c]) + (size_t)(kPad - s->koff[c]) * s->esize,
                    s->shape[c][0] * s->shape[c][1], (int)s->shape[c][2], (int)s->shape[c][2], s->pitch[c], s->esize,
                    s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(sync_exec(s));
  return GMES_OK;
}

int gmes_sim_read_field(gmes_sim* s, int32_t c, void* host, int64_t bytes) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (c < 0 || c > 5) return set_error(GMES_EINVAL, "bad component");
  if (bytes != s->ncell[c] * s->esize) return set_error(GMES_EINVAL, "size mismatch");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));  // steps queued on the sim's (or its batch's) stream come first
  if (s->pitch[c] == s->shape[c][2] && s->koff[c] == 0) {
    HIP_TRY(hipMemcpyAsync(host, static_cast<const char*>(s->f[c]) + (size_t)kPad * s->esize, (size_t)bytes,
                           hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(sync_exec(s));
    return GMES_OK;
  }
  // rows packed on the device, then one contiguous download
  int rc = stage_reserve(s, bytes);
  if (rc) return rc;
  launch_pitch_copy(static_cast<const char*>(s->f[c]) + (size_t)(kPad - s->koff[c]) * s->esize, s->d_stage,
                    s->shape[c][0] * s->shape[c][1], (int)s->shape[c][2], s->pitch[c], (int)s->shape[c][2], s->esize,
                    s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host, s->d_stage, (size_t)bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(sync_exec(s));
  return GMES_OK;
}

// ---- snapshot / restore
// The blob of a sim: a header (magic, step), the CPML/UPML and ADE state,
// the DFT accumulators, and for every attached auxiliary grid its six
// fields followed by its own blob.  The sim's own six fields are not in
// it: the caller keeps them (gmes_sim_read_field / gmes_sim_write_field).
// Restore sets the step clock back to the snapshot's and drops unread
// probe values, so the next step records and injects at the restored time.
constexpr int64_t kSnapMagic = 0x31504e5353454d47LL;  // "GMESSNP1"

static void state_blocks(gmes_sim* s, std::vector<std::pair<void*, size_t>>& out) {
  for (int c = 0; c < 6; ++c) {
    if (s->pml_state[c]) {
      const int pw = s->has_cls[c][GMES_CLS_CPML] ? 2 : 1;
      out.push_back({s->pml_state[c], (size_t)pw * s->dsize[c] * s->esize});
    }
    for (int k = GMES_CLS_DRUDE; k < GMES_NCLS; ++k)
      if (s->ade[c][k].n) out.push_back({s->ade[c][k].state, (size_t)s->ade[c][k].nstate * s->ade[c][k].n * ade_word(s, k)});
  }
  for (auto& d : s->dfts) out.push_back({d.acc, 2 * (size_t)d.nfreq * d.ncells * 2 * sizeof(double)});
}

static int64_t blob_bytes(gmes_sim* s) {
  std::vector<std::pair<void*, size_t>> b;
  state_blocks(s, b);
  int64_t t = 2 * sizeof(int64_t);
  for (auto& x : b) t += (int64_t)x.second;
  for (int i = 0; i < s->naux; ++i) {
    if (!s->aux[i]) continue;
    for (int c = 0; c < 6; ++c) t += s->aux[i]->dsize[c] * s->aux[i]->esize;
    t += blob_bytes(s->aux[i]);
  }
  return t;
}

// Copies the blob of s to or from host memory at p (the streams of s are
// idle); returns the bytes done.
static int blob_copy(gmes_sim* s, char* p, bool save, int64_t& done) {
  int64_t head[2] = {kSnapMagic, s->step};
  if (save) {
    memcpy(p, head, sizeof head);
  } else {
    memcpy(head, p, sizeof head);
    if (head[0] != kSnapMagic || head[1] < 0) return set_error(GMES_EINVAL, "not a snapshot of this engine");
  }
  done = sizeof head;
  std::vector<std::pair<void*, size_t>> b;
  state_blocks(s, b);
  for (auto& x : b) {
    if (save) HIP_TRY(hipMemcpy(p + done, x.first, x.second, hipMemcpyDeviceToHost));
    else HIP_TRY(hipMemcpy(x.first, p + done, x.second, hipMemcpyHostToDevice));
    done += (int64_t)x.second;
  }
  for (int i = 0; i < s->naux; ++i) {
    gmes_sim* a = s->aux[i];
    if (!a) continue;
    for (int c = 0; c < 6; ++c) {
      const size_t fb = (size_t)a->dsize[c] * a->esize;
      if (save) HIP_TRY(hipMemcpy(p + done, a->f[c], fb, hipMemcpyDeviceToHost));
      else HIP_TRY(hipMemcpy(a->f[c], p + done, fb, hipMemcpyHostToDevice));
      done += (int64_t)fb;
    }
    int64_t sub = 0;
    int rc = blob_copy(a, p + done, save, sub);
    if (rc) return rc;
    done += sub;
  }
  if (!save) {
    s->step = head[1];
    s->probe_base = s->step;
    s->probe_n = 0;
    std::fill(s->probe_read.begin(), s->probe_read.end(), 0);
    Ctl c;
    HIP_TRY(hipMemcpy(&c, s->d_clock, sizeof c, hipMemcpyDeviceToHost));
    c.clock = s->step;
    c.probe_base = s->probe_base;
    c.err = 0;
    HIP_TRY(hipMemcpy(s->d_clock, &c, sizeof c, hipMemcpyHostToDevice));
  }
  return GMES_OK;
}

int gmes_sim_state_bytes(const gmes_sim* s, int64_t* bytes) {
  ApiGuard api;
  if (!s || !bytes) return set_error(GMES_EINVAL, "null argument");
  *bytes = blob_bytes(const_cast<gmes_sim*>(s));
  return GMES_OK;
}

int gmes_sim_snapshot(gmes_sim* s, void* host, int64_t bytes) {
  ApiGuard api;
  if (!s || !host) return set_error(GMES_EINVAL, "null argument");
  if (blob_bytes(s) != bytes) return set_error(GMES_EINVAL, "size mismatch");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  int64_t done = 0;
  return blob_copy(s, static_cast<char*>(host), true, done);
}

int gmes_sim_restore(gmes_sim* s, const void* host, int64_t bytes) {
  ApiGuard api;
  if (!s || !host) return set_error(GMES_EINVAL, "null argument");
  if (s->parent) return set_error(GMES_EINVAL, "an auxiliary grid is restored with its parent");
  if (blob_bytes(s) != bytes) return set_error(GMES_EINVAL, "size mismatch");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  int64_t done = 0;
  return blob_copy(s, const_cast<char*>(static_cast<const char*>(host)), false, done);
}

int gmes_sim_field_sum_sq(gmes_sim* s, int32_t c, const int64_t lo[3], const int64_t hi[3], double* out) {
  ApiGuard api;
  int64_t ncells = 0;
  int rc = sq_check(s, c, lo, hi, out, ncells);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(s->device));
  const int64_t need = sq_blocks(ncells) + 1;
  if (need > s->sq_cap) {
    HIP_TRY(sync_exec(s));
    if (s->d_sq) dev_free(s, s->d_sq, (size_t)s->sq_cap * sizeof(double));
    s->d_sq = nullptr;
    s->sq_cap = 0;
    if ((rc = dev_alloc(s, (void**)&s->d_sq, (size_t)need * sizeof(double)))) return rc;
    s->sq_cap = need;
  }
  return sq_run(s, c, lo, hi, out, s->d_sq, ncells);
}

int gmes_sim_field_sum_sq_noalloc(gmes_sim* s, int32_t c, const int64_t lo[3], const int64_t hi[3], double* out) {
  ApiGuard api;
  int64_t ncells = 0;
  int rc = sq_check(s, c, lo, hi, out, ncells);
  if (rc) return rc;
  if (sq_blocks(ncells) > kScratchPartials)
    return set_error(GMES_EINVAL, "box of %lld cells exceeds the no-allocation limit of %lld cells",
                     (long long)ncells, (long long)(kScratchPartials * kSqBlockCells));
  if (s->device < 0 || s->device >= kMaxDevices) return set_error(GMES_EINVAL, "device ordinal beyond %d", kMaxDevices);
  HIP_TRY(hipSetDevice(s->device));
  std::lock_guard<std::mutex> lk(g_scratch_mu[s->device]);
  if (!g_scratch_ptr[s->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(GMES_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[s->device] = static_cast<double*>(p);
  }
  return sq_run(s, c, lo, hi, out, g_scratch_ptr[s->device], ncells);
}

int gmes_sim_step(gmes_sim* s, int64_t nsteps) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  return do_step(s, nsteps);
}

int gmes_sim_sync(gmes_sim* s) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  Ctl c;
  HIP_TRY(hipMemcpy(&c, s->d_clock, sizeof c, hipMemcpyDeviceToHost));
  if (c.err == 1) return set_error(GMES_ERUNTIME, "Dm2 corrector produced an invalid error");
  if (c.err == 2) return set_error(GMES_ERUNTIME, "Dm2 corrector failed to converge");
  for (int i = 0; i < s->naux; ++i) {
    if (!s->aux[i]) continue;
    int rc = gmes_sim_sync(s->aux[i]);
    if (rc) return rc;
  }
  if (s->timing.on) return collect_timing(s);
  return GMES_OK;
}

int gmes_sim_get_step(const gmes_sim* s, int64_t* step) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  *step = s->step;
  return GMES_OK;
}

int gmes_sim_set_graph(gmes_sim* s, int32_t enable) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  s->use_graph = enable != 0;
  destroy_graph(s);
  return GMES_OK;
}

int gmes_sim_set_ghost_fusion(gmes_sim* s, int32_t enable, int32_t* fused) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  s->allow_fuse = enable != 0;
  decide_fusion(s);
  if (fused) *fused = (s->fuse_e || s->fuse_hz) ? 1 : 0;
  return GMES_OK;
}

int gmes_sim_set_tile(gmes_sim* s, int32_t bx, int32_t by) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (bx < 0 || by < 0 || bx * by > 256 || (bx && (bx & (bx - 1))))
    return set_error(GMES_EINVAL, "bad tile: bx must be a power of two and bx*by <= 256");
  s->tile_bx = bx;
  s->tile_by = by;
  const char* ic = getenv("GMES_ICHUNK");
  s->tile_ic = ic ? atoi(ic) : 0;
  destroy_graph(s);
  return GMES_OK;
}

int gmes_sim_autotune(gmes_sim* s, int32_t* bx, int32_t* by) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  if (!s->committed) return set_error(GMES_EINVAL, "sim is not committed");
  HIP_TRY(hipSetDevice(s->device));
  int tbx = 0, tby = 0, rc;
  switch (s->g.dtype) {
    case GMES_F32: rc = autotune_t<float>(s, tbx, tby); break;
    case GMES_F64: rc = autotune_t<double>(s, tbx, tby); break;
    case GMES_C64: rc = autotune_t<cplx<float>>(s, tbx, tby); break;
    default: rc = autotune_t<cplx<double>>(s, tbx, tby); break;
  }
  if (rc) return rc;
  if (tbx > 0) {
    s->tile_bx = tbx;
    s->tile_by = tby;
    destroy_graph(s);
  }
  if (bx) *bx = tbx;
  if (by) *by = tby;
  return GMES_OK;
}

int gmes_sim_profile(gmes_sim* s, int32_t enable) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  collect_timing(s);
  s->timing.on = enable != 0;
  for (int i = 0; i < 6; ++i) { s->timing.ms[i] = 0; s->timing.launches[i] = 0; }
  return GMES_OK;
}

int gmes_sim_read_profile(gmes_sim* s, double ms[6], int64_t launches[6]) {
  ApiGuard api;
  if (!s) return set_error(GMES_EINVAL, "null sim handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(sync_exec(s));
  int rc = collect_timing(s);
  if (rc) return rc;
  for (int i = 0; i < 6; ++i) { ms[i] = s->timing.ms[i]; launches[i] = s->timing.launches[i]; }
  return GMES_OK;
}

// ----------------------------------------------------------------- batches
}  // extern "C"

struct gmes_batch {
  std::vector<gmes_sim*> sims;
  int device = 0;
  int dtype = 0;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint64_t> gens;  // member generations the graph was captured with
  // job tables: one device buffer, rebuilt on the host for every call
  void* d_jobs = nullptr;
  size_t jobs_cap = 0;
  std::vector<char> h_jobs;
  struct Launch {
    size_t off = 0;
    int njobs = 0;
    int64_t nblocks = 0;
  };
  // the launches of one step of a group of sims
  struct JobSet {
    Launch main3d[2], main2z[2], mainnd[2], src[2], ghost[2];  // [0] E, [1] H
    Launch main3d_pp[2];  // 3-D jobs with the CPML row prefetch (kBatch3dPP)
    Launch main3d_ur[2];  // 3-D jobs with the UPML row prefetch (kBatch3dUR)
    bool lds3d[2] = {false, false};  // every 3-D job stages its records in LDS
    bool lds2z[2] = {false, false};  // every collapsed-z job does
    Launch ade[GMES_NCLS];                                      // by dispersive class
    Launch dft;
  };
  JobSet set;                      // the members
  JobSet aux_set;                  // their auxiliary (TFSF) grids, stepped inside the members' step
  std::vector<gmes_sim*> aux_sims;
  std::vector<int64_t> shape_sig;  // launch shapes the graph was captured with
  // kernel timing (gmes_batch_profile): [E 3-D main, H 3-D main, E 2-D/1-D
  // main, H 2-D/1-D main, dispersive, the rest]
  gmes_engine::Timing timing;
};

namespace {

size_t append_table(std::vector<char>& h, const void* p, size_t bytes) {
  const size_t off = (h.size() + 255) / 256 * 256;
  h.resize(off + bytes);
  if (bytes) memcpy(h.data() + off, p, bytes);
  return off;
}

// Job tables of every batched launch.  The tiles follow each member's own
// setting (gmes_sim_set_tile / gmes_sim_autotune), with by = 256 / bx.
template <class T>
void build_jobs(gmes_batch* b, const std::vector<gmes_sim*>& sims, gmes_batch::JobSet& out) {
  for (int e = 0; e < 2; ++e) {
    const bool electric = e == 0;
    std::vector<MainJob<T>> j3, j2, jn, j3p, j3u;
    int64_t nb3 = 0, nb2 = 0, nbn = 0, nb3p = 0, nb3u = 0;
    bool lds3d = true, lds2z = true;
    // the block-wide record staging of the main tiles (RecStage::use_lds)
    constexpr int chunks = (int)(sizeof(DevRec<typename real_of<T>::type>) / 16);
    std::vector<SrcJob<T>> js;
    int64_t nbs = 0;
    std::vector<GhostJob<T>> jg;
    int64_t nbg = 0;
    for (gmes_sim* s : sims) {
      MainJob<T> J{};
      int ext[3];
      if (main_args<T>(s, electric, J.m, ext)) {
        int bx, by;
        tile_of(s, ext, bx, by);
        J.bx = bx;
        const int tby = 256 / bx;
        if (is3d(s)) {
          const bool fits = J.m.nrec <= kRecLds && J.m.nrec * chunks <= 256;
          J.gx = (int)(((int64_t)s->g.ny * s->g.nz + 255) / 256);
          J.gy = 1;
          if (std::is_same<T, float>::value && J.m.pp && fits) {
            J.block0 = nb3p;
            nb3p += (int64_t)J.gx * J.gy * s->g.nx;
            j3p.push_back(J);
          } else if (std::is_same<T, float>::value && J.m.ur && fits) {
            J.block0 = nb3u;
            nb3u += (int64_t)J.gx * J.gy * s->g.nx;
            j3u.push_back(J);
          } else {
            if (!fits) lds3d = false;
            J.block0 = nb3;
            nb3 += (int64_t)J.gx * J.gy * s->g.nx;
            j3.push_back(J);
          }
        } else if (is_2z(s)) {
          if (J.m.nrec > kRecLds || J.m.nrec * chunks > 256) lds2z = false;
          J.bx = ext[1];
          J.gx = (int)(((int64_t)ext[0] * ext[1] + 255) / 256);
          J.gy = 1;
          J.block0 = nb2;
          nb2 += (int64_t)J.gx * J.gy;
          j2.push_back(J);
        } else {
          J.gx = J.m.kloop ? 1 : (ext[2] + bx - 1) / bx;
          J.gy = (ext[1] + tby - 1) / tby;
          J.block0 = nbn;
          nbn += (int64_t)J.gx * J.gy * ext[0];
          jn.push_back(J);
        }
      }
      SrcJob<T> S{};
      S.a = src_args<T>(s, electric);
      if (S.a.ngroups > 0) {
        S.block0 = nbs;
        nbs += (S.a.ngroups + 255) / 256;
        js.push_back(S);
      }
      if (electric && s->fuse_e) continue;  // E ghosts written by the 3-D E kernel
      GhostJob<T> G{};
      int64_t total = 0;
      G.g = ghost_args<T>(s, electric, 0, total);
      G.block0 = nbg;
      nbg += std::max<int64_t>(1, (total + 255) / 256);  // >= 1: the block that ticks the clock
      jg.push_back(G);
    }
    out.main3d[e] = {append_table(b->h_jobs, j3.data(), j3.size() * sizeof(MainJob<T>)), (int)j3.size(), nb3};
    out.main3d_pp[e] = {append_table(b->h_jobs, j3p.data(), j3p.size() * sizeof(MainJob<T>)), (int)j3p.size(), nb3p};
    out.main3d_ur[e] = {append_table(b->h_jobs, j3u.data(), j3u.size() * sizeof(MainJob<T>)), (int)j3u.size(), nb3u};
    out.lds3d[e] = lds3d;
    out.lds2z[e] = lds2z;
    out.main2z[e] = {append_table(b->h_jobs, j2.data(), j2.size() * sizeof(MainJob<T>)), (int)j2.size(), nb2};
    out.mainnd[e] = {append_table(b->h_jobs, jn.data(), jn.size() * sizeof(MainJob<T>)), (int)jn.size(), nbn};
    out.src[e] = {append_table(b->h_jobs, js.data(), js.size() * sizeof(SrcJob<T>)), (int)js.size(), nbs};
    out.ghost[e] = {append_table(b->h_jobs, jg.data(), jg.size() * sizeof(GhostJob<T>)), (int)jg.size(), nbg};
  }
  for (int cls = GMES_CLS_DRUDE; cls < GMES_NCLS; ++cls) {
    std::vector<AdeJob<T>> ja;
    int64_t nba = 0;
    for (gmes_sim* s : sims)
      for (int c = 0; c < 3; ++c) {
        if (s->ade[c][cls].n == 0) continue;
        AdeJob<T> A{};
        A.a = ade_args<T>(s, c, cls);
        A.comp = c;
        A.block0 = nba;
        nba += (A.a.n + 255) / 256;
        ja.push_back(A);
      }
    out.ade[cls] = {append_table(b->h_jobs, ja.data(), ja.size() * sizeof(AdeJob<T>)), (int)ja.size(), nba};
  }
  std::vector<DftJob<T>> jd;
  int64_t nbd = 0;
  for (gmes_sim* s : sims) dft_jobs<T>(s, jd, nbd);
  out.dft = {append_table(b->h_jobs, jd.data(), jd.size() * sizeof(DftJob<T>)), (int)jd.size(), nbd};
}

template <class J>
const J* job_ptr(const gmes_batch* b, size_t off) {
  return reinterpret_cast<const J*>(static_cast<const char*>(b->d_jobs) + off);
}

// One step of a group of sims (the members, or their auxiliary grids)
// with one launch per kernel kind.
// HIP events around the launches of one kernel family of a batch step
// (members only; the auxiliary grids count as "the rest").
void btime_begin(gmes_batch* b, bool on, int fam, hipStream_t st) {
  if (!on || !b->timing.on) return;
  hipEvent_t x, y;
  (void)hipEventCreate(&x);
  (void)hipEventCreate(&y);
  (void)hipEventRecord(x, st);
  b->timing.pending.push_back({fam, {x, y}});
}

void btime_end(gmes_batch* b, bool on, hipStream_t st) {
  if (!on || !b->timing.on) return;
  (void)hipEventRecord(b->timing.pending.back().second.second, st);
}

template <class T>
void enqueue_set_step_t(gmes_batch* b, const gmes_batch::JobSet& js, const std::vector<gmes_sim*>& sims,
                        bool with_aux, hipStream_t st) {
  const bool tm = with_aux;  // timing covers the members' launches
  const auto& H = js.ghost[1];
  btime_begin(b, tm, 5, st);
  launch_ghost_batch<T>(job_ptr<GhostJob<T>>(b, H.off), H.njobs, H.nblocks, 1, st);
  btime_end(b, tm, st);
  for (int e = 0; e < 2; ++e) {
    const bool electric = e == 0;
    if (!electric) {
      const auto& G = js.ghost[0];
      btime_begin(b, tm, 5, st);
      launch_ghost_batch<T>(job_ptr<GhostJob<T>>(b, G.off), G.njobs, G.nblocks, 0, st);
      btime_end(b, tm, st);
    }
    btime_begin(b, tm, e, st);
    launch_main_batch<T>(electric, kBatch3dPP, job_ptr<MainJob<T>>(b, js.main3d_pp[e].off), js.main3d_pp[e].njobs,
                         js.main3d_pp[e].nblocks, st);
    launch_main_batch<T>(electric, kBatch3dUR, job_ptr<MainJob<T>>(b, js.main3d_ur[e].off), js.main3d_ur[e].njobs,
                         js.main3d_ur[e].nblocks, st);
    launch_main_batch<T>(electric, js.lds3d[e] ? kBatch3dLds : kBatch3d, job_ptr<MainJob<T>>(b, js.main3d[e].off),
                         js.main3d[e].njobs,
                         js.main3d[e].nblocks, st);
    btime_end(b, tm, st);
    btime_begin(b, tm, 2 + e, st);
    launch_main_batch<T>(electric, js.lds2z[e] ? kBatch2zLds : kBatch2z, job_ptr<MainJob<T>>(b, js.main2z[e].off),
                         js.main2z[e].njobs,
                         js.main2z[e].nblocks, st);
    launch_main_batch<T>(electric, kBatchGeneric, job_ptr<MainJob<T>>(b, js.mainnd[e].off), js.mainnd[e].njobs,
                         js.mainnd[e].nblocks, st);
    btime_end(b, tm, st);
    if (electric) {
      btime_begin(b, tm, 4, st);
      for (int cls = GMES_CLS_DRUDE; cls < GMES_NCLS; ++cls)
        launch_ade_batch<T>(cls, job_ptr<AdeJob<T>>(b, js.ade[cls].off), js.ade[cls].njobs, js.ade[cls].nblocks, st);
      btime_end(b, tm, st);
    }
    btime_begin(b, tm, 5, st);
    launch_src_batch<T>(job_ptr<SrcJob<T>>(b, js.src[e].off), js.src[e].njobs, js.src[e].nblocks, st);
    for (gmes_sim* s : sims) enqueue_probe<T>(s, electric, st);
    // every member's auxiliary grids advance one whole step here, all of
    // them in one launch per kernel kind (as each single sim does it with
    // its own grids)
    if (electric && with_aux && !b->aux_sims.empty()) enqueue_set_step_t<T>(b, b->aux_set, b->aux_sims, false, st);
    btime_end(b, tm, st);
  }
  btime_begin(b, tm, 5, st);
  launch_dft_batch<T>(job_ptr<DftJob<T>>(b, js.dft.off), js.dft.njobs, js.dft.nblocks, st);
  btime_end(b, tm, st);
}

// One step of every member, in the single-sim order, with the main, source
// and ghost kernels of all members batched into one launch each.
template <class T>
void enqueue_batch_step_t(gmes_batch* b, hipStream_t st) {
  enqueue_set_step_t<T>(b, b->set, b->sims, true, st);
}

void enqueue_batch_step(gmes_batch* b, hipStream_t st) {
  switch (b->dtype) {
    case GMES_F32: enqueue_batch_step_t<float>(b, st); break;
    case GMES_F64: enqueue_batch_step_t<double>(b, st); break;
    case GMES_C64: enqueue_batch_step_t<cplx<float>>(b, st); break;
    default: enqueue_batch_step_t<cplx<double>>(b, st); break;
  }
}

void destroy_batch_graph(gmes_batch* b) {
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
}

}  // namespace

extern "C" {

int gmes_batch_create(gmes_sim* const* sims, int32_t n, gmes_batch** out) {
  ApiGuard api;
  if (n < 1) return set_error(GMES_EINVAL, "empty batch");
  for (int i = 0; i < n; ++i) {
    gmes_sim* s = sims[i];
    if (!s) return set_error(GMES_EINVAL, "null batch member");
    if (s->device != sims[0]->device) return set_error(GMES_EINVAL, "batch members must share a device");
    if (s->g.dtype != sims[0]->g.dtype) return set_error(GMES_ETYPE, "batch members must share a storage type");
    if (s->batch_stream) return set_error(GMES_EINVAL, "sim %d is already in a batch", i);
    if (s->parent) return set_error(GMES_EINVAL, "an auxiliary grid cannot be a batch member");
    if (!s->committed) return set_error(GMES_EINVAL, "sim %d is not committed", i);
    for (int k = 0; k < i; ++k)
      if (sims[k] == s) return set_error(GMES_EINVAL, "sim %d appears twice", i);
  }
  HIP_TRY(hipSetDevice(sims[0]->device));
  gmes_batch* b = new gmes_batch();
  b->sims.assign(sims, sims + n);
  b->device = sims[0]->device;
  b->dtype = sims[0]->g.dtype;
  // non-blocking: host threads that build other sims meanwhile issue
  // null-stream copies, which must not order against (or invalidate a
  // graph capture on) the batch stream; ordering with the members is
  // explicit (sync_exec waits on the batch stream)
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return set_error(GMES_EDEVICE, "hipStreamCreate failed");
  }
  for (gmes_sim* s : b->sims) {
    (void)sync_exec(s);
    s->batch_stream = b->stream;
  }
  // setup work of the members issued on the null stream (zero fills, table
  // uploads) is not ordered before the non-blocking batch stream: drain it
  HIP_TRY(device_drain());
  *out = b;
  return GMES_OK;
}

int gmes_batch_destroy(gmes_batch* b) {
  ApiGuard api;
  if (!b) return GMES_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  destroy_batch_graph(b);
  for (auto& p : b->timing.pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (gmes_sim* s : b->sims) s->batch_stream = nullptr;
  if (b->d_jobs) (void)dev_release(b->d_jobs);
  (void)stream_release(b->stream);
  delete b;
  return GMES_OK;
}

int gmes_batch_set_graph(gmes_batch* b, int32_t enable) {
  ApiGuard api;
  if (!b) return set_error(GMES_EINVAL, "null batch handle");
  b->use_graph = enable != 0;
  destroy_batch_graph(b);
  return GMES_OK;
}

// Advance every member by nsteps: one batched launch per kernel kind and
// half step, replayed from one captured graph per step.
int gmes_batch_step(gmes_batch* b, int64_t nsteps) {
  ApiGuard api;
  if (!b) return set_error(GMES_EINVAL, "null batch handle");
  if (nsteps < 0) return set_error(GMES_EINVAL, "negative step count");
  for (gmes_sim* s : b->sims) {
    if (s->timing.on) return set_error(GMES_EINVAL, "kernel timing is per sim; switch it off for batch members");
    if (s->aux_lost) return set_error(GMES_EINVAL, "an auxiliary grid of a batch member was destroyed");
    int rc = check_windows(s, nsteps);
    if (rc) return rc;
    if ((rc = ensure_probe_capacity(s, nsteps))) return rc;
  }
  HIP_TRY(hipSetDevice(b->device));
  // the job tables are read by kernels still in flight: wait before rewriting
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->aux_sims.clear();
  for (gmes_sim* s : b->sims)
    for (int i = 0; i < s->naux; ++i)
      if (s->aux[i]) b->aux_sims.push_back(s->aux[i]);
  b->h_jobs.clear();
  switch (b->dtype) {
    case GMES_F32: build_jobs<float>(b, b->sims, b->set); build_jobs<float>(b, b->aux_sims, b->aux_set); break;
    case GMES_F64: build_jobs<double>(b, b->sims, b->set); build_jobs<double>(b, b->aux_sims, b->aux_set); break;
    case GMES_C64:
      build_jobs<cplx<float>>(b, b->sims, b->set);
      build_jobs<cplx<float>>(b, b->aux_sims, b->aux_set);
      break;
    default:
      build_jobs<cplx<double>>(b, b->sims, b->set);
      build_jobs<cplx<double>>(b, b->aux_sims, b->aux_set);
      break;
  }
  if (b->h_jobs.size() > b->jobs_cap) {
    destroy_batch_graph(b);
    if (b->d_jobs) HIP_TRY(dev_release(b->d_jobs));
    b->d_jobs = nullptr;
    b->jobs_cap = 0;
    HIP_TRY(hipMalloc(&b->d_jobs, b->h_jobs.size()));
    b->jobs_cap = b->h_jobs.size();
  }
  HIP_TRY(hipMemcpy(b->d_jobs, b->h_jobs.data(), b->h_jobs.size(), hipMemcpyHostToDevice));
  // launch shapes and per-member captured arguments decide whether the
  // graph is still valid; table contents are read at run time
  std::vector<int64_t> sig;
  for (const gmes_batch::JobSet* js : {&b->set, &b->aux_set}) {
    for (int e = 0; e < 2; ++e)
      for (const auto* L : {&js->main3d[e], &js->main3d_pp[e], &js->main3d_ur[e], &js->main2z[e], &js->mainnd[e], &js->src[e],
                            &js->ghost[e]}) {
        sig.push_back((int64_t)L->off);
        sig.push_back(L->njobs);
        sig.push_back(L->nblocks);
      }
    for (const auto& L : js->ade) {
      sig.push_back((int64_t)L.off);
      sig.push_back(L.njobs);
      sig.push_back(L.nblocks);
    }
    sig.push_back((int64_t)js->dft.off);
    sig.push_back(js->dft.njobs);
    sig.push_back(js->dft.nblocks);
  }
  std::vector<uint64_t> gens;
  for (gmes_sim* s : b->sims) {
    gens.push_back(s->gen);
    for (int i = 0; i < s->naux; ++i)
      if (s->aux[i]) gens.push_back(s->aux[i]->gen);
  }
  if (sig != b->shape_sig || gens != b->gens) destroy_batch_graph(b);
  b->shape_sig = sig;
  b->gens = gens;
  if (b->use_graph && !b->timing.on) {
    if (!b->graph_exec) {
      hipGraph_t graph;
      {
        CaptureGuard capture;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_batch_step(b, b->stream);
        HIP_TRY(hipStreamEndCapture(b->stream, &graph));
      }
      hipError_t e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return set_error(GMES_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    for (int64_t t = 0; t < nsteps; ++t) HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
  } else {
    for (int64_t t = 0; t < nsteps; ++t) enqueue_batch_step(b, b->stream);
    HIP_TRY(hipGetLastError());
  }
  for (gmes_sim* s : b->sims) advance_host_clock(s, nsteps);
  return GMES_OK;
}

int gmes_batch_profile(gmes_batch* b, int32_t enable) {
  ApiGuard api;
  if (!b) return set_error(GMES_EINVAL, "null batch handle");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (auto& p : b->timing.pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  b->timing = Timing();
  b->timing.on = enable != 0;
  return GMES_OK;
}

int gmes_batch_read_profile(gmes_batch* b, double ms[6], int64_t launches[6]) {
  ApiGuard api;
  if (!b || !ms || !launches) return set_error(GMES_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(b->device));
  for (auto& p : b->timing.pending) {
    float t = 0.f;
    HIP_TRY(hipEventSynchronize(p.second.second));
    HIP_TRY(hipEventElapsedTime(&t, p.second.first, p.second.second));
    b->timing.ms[p.first] += t;
    b->timing.launches[p.first] += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  b->timing.pending.clear();
  for (int i = 0; i < 6; ++i) {
    ms[i] = b->timing.ms[i];
    launches[i] = b->timing.launches[i];
  }
  return GMES_OK;
}

int gmes_batch_sync(gmes_batch* b) {
  ApiGuard api;
  if (!b) return set_error(GMES_EINVAL, "null batch handle");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (gmes_sim* s : b->sims) {
    int rc = gmes_sim_sync(s);
    if (rc) return rc;
  }
  return GMES_OK;
}

}  // extern "C"
