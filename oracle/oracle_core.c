/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.  CPU core shared by the
 * oracle (oracle_update.c) and the reference harness (ref_update.cc); see
 * oracle_core.h for what it restates.  FP64 / complex128 only: the
 * reference instantiates only double and std::complex<double>
 * (src/pw_material.i:386-389).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_core.h"

/* update_<c> argument table, gmes/fdtd.py:673-809:
 * Ex(hz,hy,dy,dz) Ey(hx,hz,dz,dx) Ez(hy,hx,dx,dy)
 * Hx(ez,ey,dy,dz) Hy(ex,ez,dz,dx) Hz(ey,ex,dx,dy) */
const int GMES_IN1[6] = {GMES_HZ, GMES_HX, GMES_HY, GMES_EZ, GMES_EX, GMES_EY};
const int GMES_IN2[6] = {GMES_HY, GMES_HZ, GMES_HX, GMES_EY, GMES_EZ, GMES_EX};
const int GMES_AX1[6] = {1, 2, 0, 1, 2, 0};
const int GMES_AX2[6] = {2, 0, 1, 2, 0, 1};
const int GMES_AX0[6] = {0, 1, 2, 0, 1, 2};

static __thread char g_err[512];

int ocore_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *gmes_last_error(void) { return g_err; }
int gmes_abi_version(void) { return GMES_ABI_VERSION; }
int gmes_device_count(int32_t *count) { *count = 0; return GMES_OK; }

double ocore_spacing(const gmes_sim *s, int axis) {
  return axis == 0 ? s->g.dx : axis == 1 ? s->g.dy : s->g.dz;
}

/* gmes/geometry.py:276-358 */
static void comp_shape(const gmes_grid_desc *g, int c, int64_t sh[3]) {
  if (!(g->active & (1u << c))) { sh[0] = sh[1] = sh[2] = 1; return; }
  int64_t n[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a) sh[a] = n[a] + 1;
  sh[GMES_AX0[c]] = n[GMES_AX0[c]];
}

/* field_at with the collapse rule of src/cpp23_support.hh:52-74: an
 * inactive (1,1,1) component always reads element 0. */
ocv ocore_get(const gmes_sim *s, int c, int64_t i, int64_t j, int64_t k) {
  int64_t flat = 0;
  if (s->ncell[c] > 1) flat = (i * s->shape[c][1] + j) * s->shape[c][2] + k;
  return ocore_get_flat(s, c, flat);
}

ocv ocore_get_flat(const gmes_sim *s, int c, int64_t flat) {
  ocv v;
  v.re = s->f[c][flat * s->lanes];
  v.im = s->lanes == 2 ? s->f[c][flat * 2 + 1] : 0.0;
  return v;
}

void ocore_set_flat(gmes_sim *s, int c, int64_t flat, ocv v) {
  s->f[c][flat * s->lanes] = v.re;
  if (s->lanes == 2) s->f[c][flat * 2 + 1] = v.im;
}

int gmes_sim_create(const gmes_grid_desc *desc, gmes_sim **out) {
  if (!desc || !out) return ocore_set_error(GMES_EINVAL, "null argument");
  if (desc->dtype != GMES_F64 && desc->dtype != GMES_C128)
    return ocore_set_error(GMES_ETYPE, "the oracle runs float64/complex128 only");
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1)
    return ocore_set_error(GMES_EINVAL, "grid sizes must be positive");
  gmes_sim *s = (gmes_sim *)calloc(1, sizeof *s);
  if (!s) return ocore_set_error(GMES_ENOMEM, "out of memory");
  s->g = *desc;
  s->lanes = desc->dtype == GMES_C128 ? 2 : 1;
  for (int c = 0; c < 6; ++c) {
    comp_shape(desc, c, s->shape[c]);
    s->ncell[c] = s->shape[c][0] * s->shape[c][1] * s->shape[c][2];
    s->f[c] = (double *)calloc((size_t)(s->ncell[c] * s->lanes), sizeof(double));
    if (!s->f[c]) { gmes_sim_destroy(s); return ocore_set_error(GMES_ENOMEM, "out of memory"); }
  }
  *out = s;
  return GMES_OK;
}

int gmes_sim_destroy(gmes_sim *s) {
  if (!s) return GMES_OK;
  obackend_free(s);
  for (int c = 0; c < 6; ++c) {
    free(s->f[c]); free(s->ids[c]);
    for (int k = 0; k < GMES_NCLS; ++k) {
      free(s->cls[c][k].cell); free(s->cls[c][k].rec); free(s->cls[c][k].state);
    }
  }
  for (int p = 0; p < 64; ++p)
    for (int c = 0; c < 6; ++c)
      for (int a = 0; a < 3; ++a) free(s->prof[p][c][a]);
  for (int i = 0; i < s->nprobe; ++i) free(s->probe[i].buf);
  for (int i = 0; i < s->ndft; ++i) free(s->dft[i].acc);
  free(s->dft_w);
  free(s->recs); free(s->coef); free(s->ops); free(s->waves);
  free(s);
  return GMES_OK;
}

int gmes_sim_shape(const gmes_sim *s, int32_t c, int64_t shape[3]) {
  if (c < 0 || c > 5) return ocore_set_error(GMES_EINVAL, "bad component");
  memcpy(shape, s->shape[c], sizeof(int64_t) * 3);
  return GMES_OK;
}

int gmes_sim_footprint(const gmes_sim *s, int64_t *bytes) {
  int64_t b = 0;
  for (int c = 0; c < 6; ++c) b += s->ncell[c] * s->lanes * 8;
  *bytes = b;
  return GMES_OK;
}

int gmes_sim_set_materials(gmes_sim *s, const gmes_mat_rec *recs, int32_t nrec,
                           const double *coef, int64_t ncoef) {
  if (nrec < 1 || nrec > 65535) return ocore_set_error(GMES_EINVAL, "bad record count");
  free(s->recs); free(s->coef);
  s->recs = (gmes_mat_rec *)malloc(sizeof(gmes_mat_rec) * nrec);
  memcpy(s->recs, recs, sizeof(gmes_mat_rec) * nrec);
  s->nrec = nrec;
  s->coef = (double *)malloc(sizeof(double) * (ncoef > 0 ? ncoef : 1));
  if (ncoef > 0) memcpy(s->coef, coef, sizeof(double) * ncoef);
  s->ncoef = ncoef;
  for (int r = 0; r < nrec; ++r) {
    if (recs[r].cls < 0 || recs[r].cls >= GMES_NCLS)
      return ocore_set_error(GMES_EINVAL, "record %d: bad class", r);
    if (recs[r].cls == GMES_CLS_DM2 && s->lanes == 2)
      return ocore_set_error(GMES_EINVAL, "Dm2 class supports real fields only");
  }
  return GMES_OK;
}

int gmes_sim_set_material_ids(gmes_sim *s, int32_t c, const uint16_t *ids, int64_t count) {
  if (c < 0 || c > 5) return ocore_set_error(GMES_EINVAL, "bad component");
  if (count != s->ncell[c])
    return ocore_set_error(GMES_EINVAL, "id count %lld != cells %lld", (long long)count,
                           (long long)s->ncell[c]);
  free(s->ids[c]);
  s->ids[c] = (uint16_t *)malloc(sizeof(uint16_t) * count);
  memcpy(s->ids[c], ids, sizeof(uint16_t) * count);
  return GMES_OK;
}

int gmes_sim_set_profile(gmes_sim *s, int32_t p, int32_t c, int32_t axis,
                         const double *vals, int32_t n, int32_t width) {
  if (p < 0 || p >= 64 || c < 0 || c > 5 || axis < 0 || axis > 2)
    return ocore_set_error(GMES_EINVAL, "bad profile address");
  if (n != s->shape[c][axis] || (width != 3 && width != 6))
    return ocore_set_error(GMES_EINVAL, "bad profile size");
  free(s->prof[p][c][axis]);
  s->prof[p][c][axis] = (double *)malloc(sizeof(double) * n * width);
  memcpy(s->prof[p][c][axis], vals, sizeof(double) * n * width);
  s->prof_n[p][c][axis] = n;
  s->prof_w[p][c][axis] = width;
  if (p + 1 > s->nprof) s->nprof = p + 1;
  return GMES_OK;
}

/* state words per cell of each class */
static int class_nstate(const gmes_mat_rec *r) {
  switch (r->cls) {
  case GMES_CLS_CPML: return 2;
  case GMES_CLS_UPML: return 1;
  case GMES_CLS_DRUDE: return 2 * r->n1;
  case GMES_CLS_DCP_ADE: return 1 + 2 * r->n1 + 2 * r->n2;
  case GMES_CLS_DCP_PLRC: return r->n1 + 2 * r->n2;
  case GMES_CLS_DM2: return 3 * r->n1;
  default: return 0;
  }
}

int gmes_sim_commit(gmes_sim *s) {
  if (!s->recs) return ocore_set_error(GMES_EINVAL, "materials not set");
  for (int c = 0; c < 6; ++c) {
    if (!(s->g.active & (1u << c))) continue;
    if (!s->ids[c]) return ocore_set_error(GMES_EINVAL, "ids of component %d not set", c);
    int64_t cnt[GMES_NCLS] = {0};
    int ns[GMES_NCLS] = {0};
    for (int64_t x = 0; x < s->ncell[c]; ++x) {
      uint16_t id = s->ids[c][x];
      if (id >= s->nrec) return ocore_set_error(GMES_ERANGE, "material id out of range");
      const gmes_mat_rec *r = &s->recs[id];
      cnt[r->cls]++;
      int w = class_nstate(r);
      if (w > ns[r->cls]) ns[r->cls] = w;
      if (r->cls == GMES_CLS_CPML || r->cls == GMES_CLS_UPML) {
        for (int a = 0; a < 3; ++a)
          if (!s->prof[r->prof][c][a])
            return ocore_set_error(GMES_EINVAL, "profile %d of component %d axis %d missing",
                                   r->prof, c, a);
      }
    }
    for (int k = 0; k < GMES_NCLS; ++k) {
      ocls_list *L = &s->cls[c][k];
      free(L->cell); free(L->rec); free(L->state);
      L->n = cnt[k]; L->nstate = ns[k];
      L->cell = (int64_t *)malloc(sizeof(int64_t) * (cnt[k] ? cnt[k] : 1));
      L->rec = (int32_t *)malloc(sizeof(int32_t) * (cnt[k] ? cnt[k] : 1));
      L->state = (ocv *)calloc((size_t)(cnt[k] * ns[k] + 1), sizeof(ocv));
      L->n = 0;
    }
    for (int64_t x = 0; x < s->ncell[c]; ++x) {
      uint16_t id = s->ids[c][x];
      ocls_list *L = &s->cls[c][s->recs[id].cls];
      L->cell[L->n] = x; L->rec[L->n] = id; L->n++;
    }
  }
  int rc = obackend_commit(s);
  if (rc) return rc;
  s->committed = 1;
  return GMES_OK;
}

int gmes_sim_set_sources(gmes_sim *s, const gmes_src_op *ops, int64_t nops, int32_t nwaves) {
  for (int64_t i = 0; i < nops; ++i) {
    if (ops[i].comp < 0 || ops[i].comp > 5 || ops[i].cell < 0 || ops[i].cell >= s->ncell[ops[i].comp])
      return ocore_set_error(GMES_ERANGE, "field index is out of bounds");
    if ((ops[i].kind == GMES_SRC_SET || ops[i].kind == GMES_SRC_SUB) &&
        (ops[i].wave < 0 || ops[i].wave >= nwaves))
      return ocore_set_error(GMES_EINVAL, "bad waveform row");
    if ((ops[i].kind == GMES_SRC_AUX_ADD || ops[i].kind == GMES_SRC_AUX_SUB) &&
        (ops[i].wave < -1 || ops[i].wave >= nwaves))
      return ocore_set_error(GMES_EINVAL, "bad envelope row");
  }
  free(s->ops);
  s->ops = (gmes_src_op *)malloc(sizeof(gmes_src_op) * (nops ? nops : 1));
  if (nops) memcpy(s->ops, ops, sizeof(gmes_src_op) * nops);
  s->nops = nops;
  s->nwaves = nwaves;
  return GMES_OK;
}

int gmes_sim_set_waveforms(gmes_sim *s, int64_t step0, int64_t nsteps, const double *values) {
  free(s->waves);
  size_t cnt = (size_t)(s->nwaves * nsteps * 2);
  s->waves = (double *)malloc(sizeof(double) * (cnt ? cnt : 1));
  if (cnt) memcpy(s->waves, values, sizeof(double) * cnt);
  s->wave_step0 = step0;
  s->wave_n = nsteps;
  return GMES_OK;
}

int gmes_sim_attach_aux(gmes_sim *s, gmes_sim *aux, int32_t *slot) {
  if (s->naux >= 8) return ocore_set_error(GMES_EINVAL, "too many auxiliary grids");
  if (aux->lanes != s->lanes) return ocore_set_error(GMES_ETYPE, "aux dtype mismatch");
  s->aux[s->naux] = aux;
  *slot = s->naux++;
  return GMES_OK;
}

int gmes_sim_add_probe(gmes_sim *s, int32_t c, int64_t cell, int32_t *id) {
  if (s->nprobe >= 64) return ocore_set_error(GMES_EINVAL, "too many probes");
  if (c < 0 || c > 5 || cell < 0 || cell >= s->ncell[c])
    return ocore_set_error(GMES_ERANGE, "probe index out of bounds");
  int i = s->nprobe++;
  s->probe[i].comp = c; s->probe[i].cell = cell;
  s->probe[i].n = 0; s->probe[i].cap = 1024;
  s->probe[i].buf = (double *)malloc(sizeof(double) * 2 * 1024);
  *id = i;
  return GMES_OK;
}

int gmes_sim_read_probe(gmes_sim *s, int32_t id, double *out, int64_t max, int64_t *n) {
  if (id < 0 || id >= s->nprobe) return ocore_set_error(GMES_EINVAL, "bad probe id");
  int64_t m = s->probe[id].n < max ? s->probe[id].n : max;
  memcpy(out, s->probe[id].buf, sizeof(double) * 2 * m);
  memmove(s->probe[id].buf, s->probe[id].buf + 2 * m, sizeof(double) * 2 * (s->probe[id].n - m));
  s->probe[id].n -= m;
  *n = m;
  return GMES_OK;
}

int gmes_sim_write_field(gmes_sim *s, int32_t c, const void *host, int64_t bytes) {
  if (c < 0 || c > 5) return ocore_set_error(GMES_EINVAL, "bad component");
  if (bytes != s->ncell[c] * s->lanes * 8) return ocore_set_error(GMES_EINVAL, "size mismatch");
  memcpy(s->f[c], host, (size_t)bytes);
  return GMES_OK;
}

int gmes_sim_read_field(gmes_sim *s, int32_t c, void *host, int64_t bytes) {
  if (c < 0 || c > 5) return ocore_set_error(GMES_EINVAL, "bad component");
  if (bytes != s->ncell[c] * s->lanes * 8) return ocore_set_error(GMES_EINVAL, "size mismatch");
  memcpy(host, s->f[c], (size_t)bytes);
  return GMES_OK;
}

/* Sum of |F|^2 over a box, FP64, in C order (the engine reduces in a
 * different order: agreement is to rounding). */
int gmes_sim_field_sum_sq(gmes_sim *s, int32_t c, const int64_t lo[3], const int64_t hi[3], double *out) {
  if (!s || !lo || !hi || !out) return ocore_set_error(GMES_EINVAL, "null argument");
  if (c < 0 || c > 5) return ocore_set_error(GMES_EINVAL, "bad component");
  for (int a = 0; a < 3; ++a)
    if (lo[a] < 0 || hi[a] > s->shape[c][a] || hi[a] <= lo[a])
      return ocore_set_error(GMES_ERANGE, "box out of bounds");
  double acc = 0.0;
  for (int64_t i = lo[0]; i < hi[0]; ++i)
    for (int64_t j = lo[1]; j < hi[1]; ++j)
      for (int64_t k = lo[2]; k < hi[2]; ++k) {
        ocv v = ocore_get(s, c, i, j, k);
        acc += v.re * v.re + v.im * v.im;
      }
  *out = acc;
  return GMES_OK;
}

int gmes_sim_field_sum_sq_noalloc(gmes_sim *s, int32_t c, const int64_t lo[3], const int64_t hi[3],
                                  double *out) {
  return gmes_sim_field_sum_sq(s, c, lo, hi, out);
}

/* Snapshot blob: magic, step, the per-class state lists, the DFT
 * accumulators, then per attached auxiliary grid its six fields and its
 * own blob.  Restore also sets the step clock and drops unread probe
 * values (the engine's semantics). */
#define OSNAP_MAGIC 0x31504e5353454d47LL

static int64_t oblob_bytes(const gmes_sim *s) {
  int64_t b = 16;
  for (int c = 0; c < 6; ++c)
    for (int k = 0; k < GMES_NCLS; ++k) b += s->cls[c][k].n * s->cls[c][k].nstate * 16;
  for (int d = 0; d < s->ndft; ++d) b += (int64_t)s->dft_nfreq * s->dft[d].ncells * 32;
  for (int i = 0; i < s->naux; ++i) {
    const gmes_sim *a = s->aux[i];
    if (!a) continue;
    for (int c = 0; c < 6; ++c) b += a->ncell[c] * a->lanes * 8;
    b += oblob_bytes(a);
  }
  return b;
}

static void ocopy(void *dst, const void *src, size_t n) { if (n) memcpy(dst, src, n); }

static int64_t oblob_copy(gmes_sim *s, char *p, int save) {
  int64_t head[2] = {OSNAP_MAGIC, s->step};
  if (save) memcpy(p, head, sizeof head);
  else memcpy(head, p, sizeof head);
  if (!save && (head[0] != OSNAP_MAGIC || head[1] < 0)) return -1;
  int64_t done = sizeof head;
  for (int c = 0; c < 6; ++c)
    for (int k = 0; k < GMES_NCLS; ++k) {
      size_t b = (size_t)(s->cls[c][k].n * s->cls[c][k].nstate * 16);
      if (save) ocopy(p + done, s->cls[c][k].state, b); else ocopy(s->cls[c][k].state, p + done, b);
      done += (int64_t)b;
    }
  for (int d = 0; d < s->ndft; ++d) {
    size_t b = (size_t)((int64_t)s->dft_nfreq * s->dft[d].ncells * 32);
    if (save) ocopy(p + done, s->dft[d].acc, b); else ocopy(s->dft[d].acc, p + done, b);
    done += (int64_t)b;
  }
  for (int i = 0; i < s->naux; ++i) {
    gmes_sim *a = s->aux[i];
    if (!a) continue;
    for (int c = 0; c < 6; ++c) {
      size_t b = (size_t)(a->ncell[c] * a->lanes * 8);
      if (save) ocopy(p + done, a->f[c], b); else ocopy(a->f[c], p + done, b);
      done += (int64_t)b;
    }
    int64_t sub = oblob_copy(a, p + done, save);
    if (sub < 0) return -1;
    done += sub;
  }
  if (!save) {
    s->step = head[1];
    s->err = 0;
    for (int i = 0; i < s->nprobe; ++i) s->probe[i].n = 0;
  }
  return done;
}

int gmes_sim_state_bytes(const gmes_sim *s, int64_t *bytes) {
  *bytes = oblob_bytes(s);
  return GMES_OK;
}

int gmes_sim_snapshot(gmes_sim *s, void *host, int64_t bytes) {
  if (bytes != oblob_bytes(s)) return ocore_set_error(GMES_EINVAL, "size mismatch");
  oblob_copy(s, (char *)host, 1);
  return GMES_OK;
}

int gmes_sim_restore(gmes_sim *s, const void *host, int64_t bytes) {
  if (bytes != oblob_bytes(s)) return ocore_set_error(GMES_EINVAL, "size mismatch");
  if (oblob_copy(s, (char *)host, 0) < 0)
    return ocore_set_error(GMES_EINVAL, "not a snapshot of this engine");
  return GMES_OK;
}

/* complex product (ac - bd, ad + bc), as std::complex<double> and numpy */
static ocv cmul(ocv a, ocv b) {
  ocv r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* Ghost copy of one component, gmes/fdtd.py:811-1097.  E components fill
 * their last transverse plane from plane 0 (first axis, then second);
 * H components fill plane 0 from their last plane.  The second copy
 * includes the line the first copy wrote. */
static void ghost_copy(gmes_sim *s, int c) {
  static const int order[6][2] = {{1, 2}, {2, 0}, {0, 1}, {1, 2}, {2, 0}, {0, 1}};
  const int64_t *sh = s->shape[c];
  const int electric = c < 3;
  for (int t = 0; t < 2; ++t) {
    int a = order[c][t];
    ocv ph = {s->g.phase[c][a][0], s->g.phase[c][a][1]};
    int64_t dst = electric ? sh[a] - 1 : 0, src = electric ? 0 : sh[a] - 1;
    int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
    for (int64_t u = 0; u < sh[b1]; ++u)
      for (int64_t v = 0; v < sh[b2]; ++v) {
        int64_t is[3], id[3];
        is[a] = src; is[b1] = u; is[b2] = v;
        id[a] = dst; id[b1] = u; id[b2] = v;
        int64_t fs = (is[0] * sh[1] + is[1]) * sh[2] + is[2];
        int64_t fd = (id[0] * sh[1] + id[1]) * sh[2] + id[2];
        ocv x = ocore_get_flat(s, c, fs);
        /* a real run multiplies by the integer 1 (exact) */
        ocv y = s->lanes == 2 ? cmul(ph, x) : x;
        ocore_set_flat(s, c, fd, y);
      }
  }
}

/* gmes/pw_source.py:74-87,106-119 (point sources) and :216-486 (TFSF) */
static void apply_sources(gmes_sim *s, int c, int64_t step) {
  for (int64_t q = 0; q < s->nops; ++q) {
    const gmes_src_op *op = &s->ops[q];
    if (op->comp != c) continue;
    ocv F = ocore_get_flat(s, c, op->cell);
    if (op->kind == GMES_SRC_SET || op->kind == GMES_SRC_SUB) {
      int64_t row = step - s->wave_step0;
      if (row < 0 || row >= s->wave_n) {
        s->err = ocore_set_error(GMES_ERUNTIME, "waveform window does not cover step %lld",
                                 (long long)step);
        return;
      }
      const double *w = s->waves + ((size_t)op->wave * s->wave_n + row) * 2;
      if (op->kind == GMES_SRC_SET) { F.re = w[0]; F.im = s->lanes == 2 ? w[1] : 0.0; }
      else { F.re = F.re - w[0]; F.im = F.im - (s->lanes == 2 ? w[1] : 0.0); }
    } else {
      const gmes_sim *a = s->aux[op->aux];
      ocv a0 = ocore_get_flat(a, op->aux_comp, op->i0);
      ocv a1 = ocore_get_flat(a, op->aux_comp, op->i1);
      if (op->wave >= 0) {
        /* Gaussian-beam turn-on envelope scales each sample first
           (gmes/source.py:1734-1747) */
        int64_t row = step - s->wave_step0;
        if (row < 0 || row >= s->wave_n) {
          s->err = ocore_set_error(GMES_ERUNTIME, "waveform window does not cover step %lld",
                                   (long long)step);
          return;
        }
        const double env = s->waves[((size_t)op->wave * s->wave_n + row) * 2];
        a0.re = env * a0.re; a0.im = env * a0.im;
        a1.re = env * a1.re; a1.im = env * a1.im;
      }
      ocv inc = {op->r0 * a0.re + op->r1 * a1.re, op->r0 * a0.im + op->r1 * a1.im};
      if (op->kind == GMES_SRC_AUX_ADD) { F.re = F.re + op->k * inc.re; F.im = F.im + op->k * inc.im; }
      else { F.re = F.re - op->k * inc.re; F.im = F.im - op->k * inc.im; }
    }
    ocore_set_flat(s, c, op->cell, F);
  }
}

/* Running DFT at the end of a step: E values are those of n = k + 1/2 and
 * H values those of n = k + 1, each with its own weight of the row. */
/* Compensated (Kahan) summation: *c keeps what the running sum *s lost. */
static void kahan_add(double *s, double *c, double x) {
  const double y = x - *c;
  const double t = *s + y;
  *c = (t - *s) - y;
  *s = t;
}

static void accumulate_dft(gmes_sim *s) {
  if (s->ndft == 0) return;
  const int64_t row = s->step - s->dft_step0;
  if (row < 0 || row >= s->dft_n) {
    s->err = ocore_set_error(GMES_ERUNTIME, "DFT weights do not cover step %lld", (long long)s->step);
    return;
  }
  for (int d = 0; d < s->ndft; ++d) {
    if (s->step % s->dft[d].decim) continue;
    const int c = s->dft[d].comp;
    const double *w = s->dft_w + row * s->dft_nfreq * 4 + (c < 3 ? 0 : 2);
    for (int64_t q = 0; q < s->dft[d].ncells; ++q) {
      const int64_t k = q % s->dft[d].n[2], j = (q / s->dft[d].n[2]) % s->dft[d].n[1];
      const int64_t i = q / (s->dft[d].n[2] * s->dft[d].n[1]);
      const ocv v = ocore_get(s, c, i + s->dft[d].lo[0], j + s->dft[d].lo[1], k + s->dft[d].lo[2]);
      for (int f = 0; f < s->dft_nfreq; ++f) {
        double *a = s->dft[d].acc + 2 * (f * s->dft[d].ncells + q);
        double *e = a + 2 * (int64_t)s->dft_nfreq * s->dft[d].ncells;  /* compensations */
        const double wr = w[4 * f], wi = w[4 * f + 1];
        kahan_add(&a[0], &e[0], wr * v.re - wi * v.im);
        kahan_add(&a[1], &e[1], wr * v.im + wi * v.re);
      }
    }
  }
}

static void record_probes(gmes_sim *s, int electric) {
  for (int i = 0; i < s->nprobe; ++i) {
    if ((s->probe[i].comp < 3) != electric) continue;
    if (s->probe[i].n == s->probe[i].cap) {
      s->probe[i].cap *= 2;
      s->probe[i].buf = (double *)realloc(s->probe[i].buf, sizeof(double) * 2 * s->probe[i].cap);
    }
    ocv v = ocore_get_flat(s, s->probe[i].comp, s->probe[i].cell);
    s->probe[i].buf[2 * s->probe[i].n] = v.re;
    s->probe[i].buf[2 * s->probe[i].n + 1] = v.im;
    s->probe[i].n++;
  }
}

/* One leapfrog step, gmes/fdtd.py:1099-1122. */
static int one_step(gmes_sim *s) {
  const double n_e = (double)s->step + 0.5, n_h = (double)s->step + 1.0;
  for (int c = 3; c < 6; ++c)
    if (s->g.active & (1u << c)) ghost_copy(s, c);
  for (int c = 0; c < 3; ++c) {
    if (!(s->g.active & (1u << c))) continue;
    int rc = obackend_update(s, c, n_e);
    if (rc) return rc;
    apply_sources(s, c, s->step);
    if (s->err) return s->err;
  }
  record_probes(s, 1);
  for (int i = 0; i < s->naux; ++i) {
    int rc = gmes_sim_step(s->aux[i], 1);
    if (rc) return rc;
  }
  for (int c = 0; c < 3; ++c)
    if (s->g.active & (1u << c)) ghost_copy(s, c);
  for (int c = 3; c < 6; ++c) {
    if (!(s->g.active & (1u << c))) continue;
    int rc = obackend_update(s, c, n_h);
    if (rc) return rc;
    apply_sources(s, c, s->step);
    if (s->err) return s->err;
  }
  record_probes(s, 0);
  accumulate_dft(s);
  if (s->err) return s->err;
  s->step++;
  return GMES_OK;
}

int gmes_sim_add_dft(gmes_sim *s, int32_t c, const int64_t lo[3], const int64_t hi[3], int32_t nfreq,
                     int32_t decim, int32_t *id) {
  if (c < 0 || c > 5 || !(s->g.active & (1u << c))) return ocore_set_error(GMES_EINVAL, "DFT of an inactive component");
  if (nfreq < 1 || decim < 1) return ocore_set_error(GMES_EINVAL, "nfreq and decim must be positive");
  if (s->ndft && nfreq != s->dft_nfreq) return ocore_set_error(GMES_EINVAL, "all DFT monitors share one frequency list");
  if (s->ndft >= 16) return ocore_set_error(GMES_EINVAL, "too many DFT monitors");
  int d = s->ndft;
  s->dft[d].comp = c;
  s->dft[d].ncells = 1;
  for (int a = 0; a < 3; ++a) {
    if (lo[a] < 0 || hi[a] > s->shape[c][a] || hi[a] <= lo[a]) return ocore_set_error(GMES_ERANGE, "DFT box out of bounds");
    s->dft[d].lo[a] = lo[a];
    s->dft[d].n[a] = hi[a] - lo[a];
    s->dft[d].ncells *= s->dft[d].n[a];
  }
  s->dft[d].decim = decim;
  s->dft[d].acc = (double *)calloc((size_t)(2 * nfreq * s->dft[d].ncells * 2), sizeof(double));
  s->dft_nfreq = nfreq;
  s->ndft++;
  *id = d;
  return GMES_OK;
}

int gmes_sim_set_dft_weights(gmes_sim *s, int64_t step0, int64_t nsteps, const double *w) {
  if (!s->ndft) return ocore_set_error(GMES_EINVAL, "no DFT monitor");
  free(s->dft_w);
  size_t cnt = (size_t)(nsteps * s->dft_nfreq * 4);
  s->dft_w = (double *)malloc(sizeof(double) * (cnt ? cnt : 1));
  if (cnt) memcpy(s->dft_w, w, sizeof(double) * cnt);
  s->dft_step0 = step0;
  s->dft_n = nsteps;
  return GMES_OK;
}

int gmes_sim_read_dft(gmes_sim *s, int32_t id, double *out, int64_t count) {
  if (id < 0 || id >= s->ndft) return ocore_set_error(GMES_EINVAL, "bad DFT id");
  if (count != s->dft_nfreq * s->dft[id].ncells * 2) return ocore_set_error(GMES_EINVAL, "size mismatch");
  memcpy(out, s->dft[id].acc, sizeof(double) * count);
  return GMES_OK;
}

int gmes_sim_step(gmes_sim *s, int64_t nsteps) {
  if (!s->committed) return ocore_set_error(GMES_EINVAL, "sim not committed");
  if (nsteps < 0) return ocore_set_error(GMES_EINVAL, "negative step count");
  if (s->err) return s->err;
  for (int64_t t = 0; t < nsteps; ++t) {
    int rc = one_step(s);
    if (rc) { s->err = rc; snprintf(s->errmsg, sizeof s->errmsg, "%s", gmes_last_error()); return rc; }
  }
  return GMES_OK;
}

int gmes_sim_sync(gmes_sim *s) {
  if (s->err) return ocore_set_error(s->err, "%s", s->errmsg[0] ? s->errmsg : gmes_last_error());
  return GMES_OK;
}

int gmes_sim_get_step(const gmes_sim *s, int64_t *step) { *step = s->step; return GMES_OK; }
int gmes_sim_set_graph(gmes_sim *s, int32_t enable) { (void)s; (void)enable; return GMES_OK; }
int gmes_sim_set_tile(gmes_sim *s, int32_t bx, int32_t by) { (void)s; (void)bx; (void)by; return GMES_OK; }
/* no launch shapes on the CPU: the tile is reported as 0 x 0 */
int gmes_sim_autotune(gmes_sim *s, int32_t *bx, int32_t *by) {
  (void)s;
  if (bx) *bx = 0;
  if (by) *by = 0;
  return GMES_OK;
}
/* the oracle always makes the separate ghost copy */
int gmes_sim_set_ghost_fusion(gmes_sim *s, int32_t enable, int32_t *fused) {
  (void)s;
  (void)enable;
  if (fused) *fused = 0;
  return GMES_OK;
}
int gmes_sim_profile(gmes_sim *s, int32_t enable) { (void)s; (void)enable; return GMES_OK; }
int gmes_sim_read_profile(gmes_sim *s, double ms[6], int64_t launches[6]) {
  (void)s;
  for (int i = 0; i < 6; ++i) { ms[i] = 0; launches[i] = 0; }
  return GMES_OK;
}

struct gmes_batch { gmes_sim **sims; int32_t n; };

int gmes_batch_create(gmes_sim *const *sims, int32_t n, gmes_batch **out) {
  gmes_batch *b = (gmes_batch *)calloc(1, sizeof *b);
  b->sims = (gmes_sim **)malloc(sizeof(gmes_sim *) * (n ? n : 1));
  memcpy(b->sims, sims, sizeof(gmes_sim *) * n);
  b->n = n;
  *out = b;
  return GMES_OK;
}
int gmes_batch_destroy(gmes_batch *b) { if (b) { free(b->sims); free(b); } return GMES_OK; }
/* Every member is stepped; a member whose update fails (DM2) keeps its
 * error and stops, and the error is reported by its sync, as the engine
 * defers device errors per member to gmes_sim_sync / gmes_batch_sync. */
int gmes_batch_step(gmes_batch *b, int64_t nsteps) {
  if (nsteps < 0) return ocore_set_error(GMES_EINVAL, "negative step count");
  for (int i = 0; i < b->n; ++i) {
    gmes_sim *s = b->sims[i];
    if (!s->committed) return ocore_set_error(GMES_EINVAL, "sim not committed");
  }
  for (int i = 0; i < b->n; ++i) {
    gmes_sim *s = b->sims[i];
    int rc = gmes_sim_step(s, nsteps);
    if (rc && !s->err) return rc;
  }
  return GMES_OK;
}
int gmes_batch_sync(gmes_batch *b) {
  for (int i = 0; i < b->n; ++i) { int rc = gmes_sim_sync(b->sims[i]); if (rc) return rc; }
  return GMES_OK;
}
int gmes_batch_set_graph(gmes_batch *b, int32_t enable) { (void)b; (void)enable; return GMES_OK; }
int gmes_batch_profile(gmes_batch *b, int32_t enable) { (void)b; (void)enable; return GMES_OK; }
int gmes_batch_read_profile(gmes_batch *b, double ms[6], int64_t launches[6]) {
  (void)b;
  for (int i = 0; i < 6; ++i) { ms[i] = 0; launches[i] = 0; }
  return GMES_OK;
}
