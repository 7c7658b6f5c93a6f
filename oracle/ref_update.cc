/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * Backend of oracle/_ref/libgmes_ref.so: drives the reference's own
 * pointwise updaters (the PwMaterial<T> plugin classes of
 * /root/reference/src/pw_*.hh, compiled unchanged from where they lie)
 * behind the same C-ABI as the engine.  For every (component, class) the
 * cells of a sim are attached to one reference updater object with the
 * parameter struct the reference's factories would build
 * (src/material.pyx get_pw_material_<c>), and obackend_update() calls
 * update_all() through the virtual base, as FDTD.update_<c> does
 * (gmes/fdtd.py:673-809).  Records of class DRUDE with n2 == 1 are
 * Lorentz media and go to the reference's LorentzE* updaters
 * (src/pw_lorentz.hh), the others to DrudeE* (src/pw_drude.hh).
 */
#include <complex>
#include <memory>
#include <stdexcept>
#include <vector>

#include "pw_const.hh"
#include "pw_cpml.hh"
#include "pw_dcp.hh"
#include "pw_dielectric.hh"
#include "pw_dm2.hh"
#include "pw_drude.hh"
#include "pw_lorentz.hh"
#include "pw_upml.hh"

#include "oracle_core.h"

namespace {

template <typename T>
struct RefSet {
  std::unique_ptr<gmes::PwMaterial<T>> up[6][GMES_NCLS];
  std::unique_ptr<gmes::PwMaterial<T>> lorentz[6];  // DRUDE records with n2 == 1
};

template <typename T>
T make_value(const double v[2]);
template <>
double make_value<double>(const double v[2]) { return v[0]; }
template <>
std::complex<double> make_value<std::complex<double>>(const double v[2]) {
  return std::complex<double>(v[0], v[1]);
}

/* per-component class constructors */
template <typename T>
gmes::PwMaterial<T>* new_updater(int cls, int c) {
  using namespace gmes;
  switch (cls) {
  case GMES_CLS_CONST:
    switch (c) { case 0: return new ConstEx<T>(); case 1: return new ConstEy<T>(); case 2: return new ConstEz<T>();
                 case 3: return new ConstHx<T>(); case 4: return new ConstHy<T>(); default: return new ConstHz<T>(); }
  case GMES_CLS_DIELECTRIC:
    switch (c) { case 0: return new DielectricEx<T>(); case 1: return new DielectricEy<T>(); case 2: return new DielectricEz<T>();
                 case 3: return new DielectricHx<T>(); case 4: return new DielectricHy<T>(); default: return new DielectricHz<T>(); }
  case GMES_CLS_CPML:
    switch (c) { case 0: return new CpmlEx<T>(); case 1: return new CpmlEy<T>(); case 2: return new CpmlEz<T>();
                 case 3: return new CpmlHx<T>(); case 4: return new CpmlHy<T>(); default: return new CpmlHz<T>(); }
  case GMES_CLS_UPML:
    switch (c) { case 0: return new UpmlEx<T>(); case 1: return new UpmlEy<T>(); case 2: return new UpmlEz<T>();
                 case 3: return new UpmlHx<T>(); case 4: return new UpmlHy<T>(); default: return new UpmlHz<T>(); }
  case GMES_CLS_DRUDE:
    switch (c) { case 0: return new DrudeEx<T>(); case 1: return new DrudeEy<T>(); case 2: return new DrudeEz<T>();
                 default: return nullptr; }
  case GMES_CLS_DCP_ADE:
    switch (c) { case 0: return new DcpAdeEx<T>(); case 1: return new DcpAdeEy<T>(); case 2: return new DcpAdeEz<T>();
                 default: return nullptr; }
  case GMES_CLS_DCP_PLRC:
    switch (c) { case 0: return new DcpPlrcEx<T>(); case 1: return new DcpPlrcEy<T>(); case 2: return new DcpPlrcEz<T>();
                 default: return nullptr; }
  default:
    return nullptr;
  }
}

template <typename T>
gmes::PwMaterial<T>* new_lorentz(int c) {
  using namespace gmes;
  switch (c) { case 0: return new LorentzEx<T>(); case 1: return new LorentzEy<T>(); case 2: return new LorentzEz<T>();
               default: return nullptr; }
}

template <typename T>
gmes::PwMaterial<T>* new_dm2(int c) {
  return nullptr;
}
template <>
gmes::PwMaterial<double>* new_dm2<double>(int c) {
  switch (c) { case 0: return new gmes::Dm2Ex<double>(); case 1: return new gmes::Dm2Ey<double>();
               case 2: return new gmes::Dm2Ez<double>(); default: return nullptr; }
}

template <typename T>
void attach_cell(gmes_sim* s, gmes::PwMaterial<T>* up, int cls, int c, int64_t x,
                 const gmes_mat_rec& r) {
  using namespace gmes;
  const int64_t sy = s->shape[c][1], sz = s->shape[c][2];
  int idx[3] = {(int)(x / (sy * sz)), (int)((x / sz) % sy), (int)(x % sz)};
  const double* cf = s->coef + r.coef_off;
  const bool electric = c < 3;
  switch (cls) {
  case GMES_CLS_CONST:
    if (electric) { ConstElectricParam<T> p; p.eps_inf = r.eps_mu; p.value = make_value<T>(r.value); up->attach(idx, 3, &p); }
    else { ConstMagneticParam<T> p; p.mu_inf = r.eps_mu; p.value = make_value<T>(r.value); up->attach(idx, 3, &p); }
    break;
  case GMES_CLS_DIELECTRIC:
    if (electric) { DielectricElectricParam<T> p; p.eps_inf = r.eps_mu; up->attach(idx, 3, &p); }
    else { DielectricMagneticParam<T> p; p.mu_inf = r.eps_mu; up->attach(idx, 3, &p); }
    break;
  case GMES_CLS_CPML: {
    const double* p1 = s->prof[r.prof][c][GMES_AX1[c]] + idx[GMES_AX1[c]] * 3;
    const double* p2 = s->prof[r.prof][c][GMES_AX2[c]] + idx[GMES_AX2[c]] * 3;
    if (electric) {
      CpmlElectricParam<T> p; p.eps_inf = r.eps_mu;
      p.b1 = p1[0]; p.c1 = p1[1]; p.kappa1 = p1[2]; p.b2 = p2[0]; p.c2 = p2[1]; p.kappa2 = p2[2];
      p.psi1 = T(0); p.psi2 = T(0);
      up->attach(idx, 3, &p);
    } else {
      CpmlMagneticParam<T> p; p.mu_inf = r.eps_mu;
      p.b1 = p1[0]; p.c1 = p1[1]; p.kappa1 = p1[2]; p.b2 = p2[0]; p.c2 = p2[1]; p.kappa2 = p2[2];
      p.psi1 = T(0); p.psi2 = T(0);
      up->attach(idx, 3, &p);
    }
    break;
  }
  case GMES_CLS_UPML: {
    const double* p1 = s->prof[r.prof][c][GMES_AX1[c]] + idx[GMES_AX1[c]] * 6;
    const double* p2 = s->prof[r.prof][c][GMES_AX2[c]] + idx[GMES_AX2[c]] * 6;
    const double* p0 = s->prof[r.prof][c][GMES_AX0[c]] + idx[GMES_AX0[c]] * 6;
    if (electric) {
      UpmlElectricParam<T> p; p.eps_inf = r.eps_mu;
      p.c1 = p1[0]; p.c2 = p1[1]; p.c3 = p2[2]; p.c4 = p2[3]; p.c5 = p0[4]; p.c6 = p0[5]; p.d = T(0);
      up->attach(idx, 3, &p);
    } else {
      UpmlMagneticParam<T> p; p.mu_inf = r.eps_mu;
      p.c1 = p1[0]; p.c2 = p1[1]; p.c3 = p2[2]; p.c4 = p2[3]; p.c5 = p0[4]; p.c6 = p0[5]; p.b = T(0);
      up->attach(idx, 3, &p);
    }
    break;
  }
  case GMES_CLS_DRUDE: {
    if (r.n2 == 1) {
      LorentzElectricParam<T> p; p.eps_inf = r.eps_mu;
      for (int q = 0; q < r.n1; ++q) p.a.push_back({cf[3 * q], cf[3 * q + 1], cf[3 * q + 2]});
      p.c = {cf[3 * r.n1], cf[3 * r.n1 + 1], cf[3 * r.n1 + 2]};
      p.l_now.assign(r.n1, T(0)); p.l_new.assign(r.n1, T(0));
      up->attach(idx, 3, &p);
      break;
    }
    DrudeElectricParam<T> p; p.eps_inf = r.eps_mu;
    for (int q = 0; q < r.n1; ++q) p.a.push_back({cf[3 * q], cf[3 * q + 1], cf[3 * q + 2]});
    p.c = {cf[3 * r.n1], cf[3 * r.n1 + 1], cf[3 * r.n1 + 2]};
    p.q_now.assign(r.n1, T(0)); p.q_new.assign(r.n1, T(0));
    up->attach(idx, 3, &p);
    break;
  }
  case GMES_CLS_DCP_ADE: {
    DcpAdeElectricParam<T> p; p.eps_inf = r.eps_mu;
    const double* b = cf + 3 * r.n1;
    const double* cc = b + 5 * r.n2;
    for (int q = 0; q < r.n1; ++q) p.a.push_back({cf[3 * q], cf[3 * q + 1], cf[3 * q + 2]});
    for (int q = 0; q < r.n2; ++q) p.b.push_back({b[5 * q], b[5 * q + 1], b[5 * q + 2], b[5 * q + 3], b[5 * q + 4]});
    p.c = {cc[0], cc[1], cc[2], cc[3]};
    p.e_old = T(0);
    p.q_old.assign(r.n1, T(0)); p.q_now.assign(r.n1, T(0));
    p.p_old.assign(r.n2, T(0)); p.p_now.assign(r.n2, T(0));
    up->attach(idx, 3, &p);
    break;
  }
  case GMES_CLS_DCP_PLRC: {
    DcpPlrcElectricParam<T> p; p.eps_inf = r.eps_mu;
    const double* b = cf + 3 * r.n1;
    const double* cc = b + 6 * r.n2;
    for (int q = 0; q < r.n1; ++q) p.a.push_back({cf[3 * q], cf[3 * q + 1], cf[3 * q + 2]});
    for (int q = 0; q < r.n2; ++q)
      p.b.push_back({std::complex<double>(b[6 * q], b[6 * q + 1]), std::complex<double>(b[6 * q + 2], b[6 * q + 3]),
                     std::complex<double>(b[6 * q + 4], b[6 * q + 5])});
    p.c = {cc[0], cc[1], cc[2]};
    p.psi_dp_re.assign(r.n1, 0.0); p.psi_dp_im.assign(r.n1, 0.0);
    p.psi_cp_re.assign(r.n2, std::complex<double>(0)); p.psi_cp_im.assign(r.n2, std::complex<double>(0));
    up->attach(idx, 3, &p);
    break;
  }
  case GMES_CLS_DM2: {
    if constexpr (std::is_same_v<T, double>) {
      Dm2ElectricParam<double> p; p.eps_inf = r.eps_mu;
      for (int q = 0; q < r.n1; ++q) {
        p.omega.push_back(cf[q]); p.n_atom.push_back(cf[r.n1 + q]);
        std::array<double, 3> u0; u0.fill(0.0); p.u.push_back(u0);
      }
      p.rho30 = cf[2 * r.n1]; p.gamma = cf[2 * r.n1 + 1]; p.t1 = cf[2 * r.n1 + 2];
      p.t2 = cf[2 * r.n1 + 3]; p.hbar = cf[2 * r.n1 + 4]; p.rtol = cf[2 * r.n1 + 5];
      up->attach(idx, 3, &p);
    }
    break;
  }
  default:
    break;
  }
}

template <typename T>
int commit_t(gmes_sim* s) {
  auto* set = new RefSet<T>();
  s->impl = set;
  for (int c = 0; c < 6; ++c) {
    if (!(s->g.active & (1u << c))) continue;
    for (int cls = 1; cls < GMES_NCLS; ++cls) {
      ocls_list* L = &s->cls[c][cls];
      if (L->n == 0) continue;
      gmes::PwMaterial<T>* up = cls == GMES_CLS_DM2 ? new_dm2<T>(c) : new_updater<T>(cls, c);
      if (!up) return ocore_set_error(GMES_EINVAL, "class %d has no updater for component %d", cls, c);
      set->up[c][cls].reset(up);
      if (cls == GMES_CLS_DRUDE) set->lorentz[c].reset(new_lorentz<T>(c));
      for (int64_t q = 0; q < L->n; ++q) {
        const gmes_mat_rec& r = s->recs[L->rec[q]];
        gmes::PwMaterial<T>* target = (cls == GMES_CLS_DRUDE && r.n2 == 1) ? set->lorentz[c].get() : up;
        attach_cell<T>(s, target, cls, c, L->cell[q], r);
      }
    }
  }
  return GMES_OK;
}

template <typename T>
int update_t(gmes_sim* s, int c, double n) {
  auto* set = static_cast<RefSet<T>*>(s->impl);
  T* F = reinterpret_cast<T*>(s->f[c]);
  const int in1 = GMES_IN1[c], in2 = GMES_IN2[c];
  const T* A = reinterpret_cast<const T*>(s->f[in1]);
  const T* B = reinterpret_cast<const T*>(s->f[in2]);
  const double d1 = ocore_spacing(s, GMES_AX1[c]), d2 = ocore_spacing(s, GMES_AX2[c]);
  try {
    for (int cls = 1; cls <= GMES_NCLS; ++cls) {
      gmes::PwMaterial<T>* up = cls < GMES_NCLS ? set->up[c][cls].get() : set->lorentz[c].get();
      if (!up || up->idx_size() == 0) continue;
      up->update_all(F, (int)s->shape[c][0], (int)s->shape[c][1], (int)s->shape[c][2],
                     A, (int)s->shape[in1][0], (int)s->shape[in1][1], (int)s->shape[in1][2],
                     B, (int)s->shape[in2][0], (int)s->shape[in2][1], (int)s->shape[in2][2],
                     d1, d2, s->g.dt, n);
    }
  } catch (const std::exception& e) {
    s->err = ocore_set_error(GMES_ERUNTIME, "%s", e.what());
    return s->err;
  }
  return GMES_OK;
}

}  // namespace

extern "C" int obackend_commit(gmes_sim* s) {
  try {
    return s->lanes == 2 ? commit_t<std::complex<double>>(s) : commit_t<double>(s);
  } catch (const std::exception& e) {
    return ocore_set_error(GMES_ERUNTIME, "%s", e.what());
  }
}

extern "C" int obackend_update(gmes_sim* s, int c, double n) {
  return s->lanes == 2 ? update_t<std::complex<double>>(s, c, n) : update_t<double>(s, c, n);
}

extern "C" void obackend_free(gmes_sim* s) {
  if (!s->impl) return;
  if (s->lanes == 2) delete static_cast<RefSet<std::complex<double>>*>(s->impl);
  else delete static_cast<RefSet<double>*>(s->impl);
  s->impl = nullptr;
}
