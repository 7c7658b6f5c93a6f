/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * The oracle's per-cell arithmetic: a plain-C restatement of the
 * reference's pointwise material updaters, one function per class, each
 * keeping the reference's operation order (divide by d1/d2, `dt/eps`
 * first, left-to-right sums) so that FP64 results are bit-identical.
 * Complex values are (re, im) pairs; every real coefficient acts on both
 * parts, complex coefficients use (ac-bd, ad+bc).  A real run keeps im = 0.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle_core.h"

typedef ocv V;
static V vadd(V a, V b) { V r = {a.re + b.re, a.im + b.im}; return r; }
static V vsub(V a, V b) { V r = {a.re - b.re, a.im - b.im}; return r; }
static V vscl(double s, V a) { V r = {s * a.re, s * a.im}; return r; }
static V vdiv(V a, double d) { V r = {a.re / d, a.im / d}; return r; }
static V cmul(V a, V b) { V r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }

/* Curl differences D1, D2 of component c at (i,j,k): the field_at reads of
 * src/pw_dielectric.hh:118-119 (Ex), :157-158 (Ey), :197-198 (Ez),
 * :292-293 (Hx), :331-332 (Hy), :370-371 (Hz). */
static void curl_terms(const gmes_sim *s, int c, int64_t i, int64_t j, int64_t k, V *D1, V *D2) {
  switch (c) {
  case GMES_EX:
    *D1 = vsub(ocore_get(s, GMES_HZ, i + 1, j + 1, k), ocore_get(s, GMES_HZ, i + 1, j, k));
    *D2 = vsub(ocore_get(s, GMES_HY, i + 1, j, k + 1), ocore_get(s, GMES_HY, i + 1, j, k));
    break;
  case GMES_EY:
    *D1 = vsub(ocore_get(s, GMES_HX, i, j + 1, k + 1), ocore_get(s, GMES_HX, i, j + 1, k));
    *D2 = vsub(ocore_get(s, GMES_HZ, i + 1, j + 1, k), ocore_get(s, GMES_HZ, i, j + 1, k));
    break;
  case GMES_EZ:
    *D1 = vsub(ocore_get(s, GMES_HY, i + 1, j, k + 1), ocore_get(s, GMES_HY, i, j, k + 1));
    *D2 = vsub(ocore_get(s, GMES_HX, i, j + 1, k + 1), ocore_get(s, GMES_HX, i, j, k + 1));
    break;
  case GMES_HX:
    *D1 = vsub(ocore_get(s, GMES_EZ, i, j, k - 1), ocore_get(s, GMES_EZ, i, j - 1, k - 1));
    *D2 = vsub(ocore_get(s, GMES_EY, i, j - 1, k), ocore_get(s, GMES_EY, i, j - 1, k - 1));
    break;
  case GMES_HY:
    *D1 = vsub(ocore_get(s, GMES_EX, i - 1, j, k), ocore_get(s, GMES_EX, i - 1, j, k - 1));
    *D2 = vsub(ocore_get(s, GMES_EZ, i, j, k - 1), ocore_get(s, GMES_EZ, i - 1, j, k - 1));
    break;
  default: /* GMES_HZ */
    *D1 = vsub(ocore_get(s, GMES_EY, i, j - 1, k), ocore_get(s, GMES_EY, i - 1, j - 1, k));
    *D2 = vsub(ocore_get(s, GMES_EX, i - 1, j, k), ocore_get(s, GMES_EX, i - 1, j - 1, k));
    break;
  }
}

static const double *prow(const gmes_sim *s, int p, int c, int axis, int64_t idx) {
  return s->prof[p][c][axis] + idx * s->prof_w[p][c][axis];
}

/* DM2 relative error, src/pw_dm2.hh:40-89 */
static double dm2_l2(double e, const double *u, int n) {
  double accu = e * e;
  for (int i = 0; i < n; ++i) {
    double ip = 0.0;
    for (int j = 0; j < 3; ++j) ip = ip + u[3 * i + j] * u[3 * i + j];
    accu += ip;
  }
  return sqrt(accu);
}

static double dm2_rel_error(double e_new, const double *u_new, double e_ref, const double *u_ref,
                            int n, double *scratch) {
  for (int i = 0; i < 3 * n; ++i) scratch[i] = u_new[i] - u_ref[i];
  double dn = dm2_l2(e_new - e_ref, scratch, n);
  double rn = dm2_l2(e_ref, u_ref, n);
  if (rn == 0) return dn == 0 ? 0 : INFINITY;
  return dn / rn;
}

int obackend_commit(gmes_sim *s) { (void)s; return GMES_OK; }
void obackend_free(gmes_sim *s) { (void)s; }

int obackend_update(gmes_sim *s, int c, double n) {
  const int electric = c < 3;
  const double d1 = ocore_spacing(s, GMES_AX1[c]), d2 = ocore_spacing(s, GMES_AX2[c]);
  const double dt = s->g.dt;
  const int64_t sy = s->shape[c][1], sz = s->shape[c][2];
  for (int cl = 0; cl < GMES_NCLS; ++cl) {
    ocls_list *L = &s->cls[c][cl];
    if (cl == GMES_CLS_DUMMY) continue;   /* src/pw_dummy.hh:62-72 */
    for (int64_t q = 0; q < L->n; ++q) {
      const int64_t x = L->cell[q];
      const int64_t i = x / (sy * sz), j = (x / sz) % sy, k = x % sz;
      const int64_t ix[3] = {i, j, k};
      const gmes_mat_rec *r = &s->recs[L->rec[q]];
      const double *cf = s->coef + r->coef_off;
      V *st = L->state + q * L->nstate;
      V F = ocore_get_flat(s, c, x);
      V D1, D2;
      if (cl == GMES_CLS_CONST) {          /* src/pw_const.hh:96,198 */
        V val = {r->value[0], s->lanes == 2 ? r->value[1] : 0.0};
        ocore_set_flat(s, c, x, val);
        continue;
      }
      curl_terms(s, c, i, j, k, &D1, &D2);
      switch (cl) {
      case GMES_CLS_DIELECTRIC:           /* src/pw_dielectric.hh:118,292 */
        if (electric) F = vadd(F, vscl(dt / r->eps_mu, vsub(vdiv(D1, d1), vdiv(D2, d2))));
        else F = vadd(F, vscl(dt / r->eps_mu, vsub(vdiv(D2, d2), vdiv(D1, d1))));
        break;
      case GMES_CLS_CPML: {               /* src/pw_cpml.hh:127-132,339-344 */
        const double *p1 = prow(s, r->prof, c, GMES_AX1[c], ix[GMES_AX1[c]]);
        const double *p2 = prow(s, r->prof, c, GMES_AX2[c], ix[GMES_AX2[c]]);
        st[0] = vadd(vscl(p1[0], st[0]), vdiv(vscl(p1[1], D1), d1));
        st[1] = vadd(vscl(p2[0], st[1]), vdiv(vscl(p2[1], D2), d2));
        V t = vsub(vadd(vsub(vdiv(vdiv(D1, d1), p1[2]), vdiv(vdiv(D2, d2), p2[2])), st[0]), st[1]);
        if (electric) F = vadd(F, vscl(dt / r->eps_mu, t));
        else F = vsub(F, vscl(dt / r->eps_mu, t));
        break;
      }
      case GMES_CLS_UPML: {               /* src/pw_upml.hh:124-128,328-332 */
        const double *p1 = prow(s, r->prof, c, GMES_AX1[c], ix[GMES_AX1[c]]);
        const double *p2 = prow(s, r->prof, c, GMES_AX2[c], ix[GMES_AX2[c]]);
        const double *p0 = prow(s, r->prof, c, GMES_AX0[c], ix[GMES_AX0[c]]);
        V dstore = st[0];
        V cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
        if (electric) st[0] = vadd(vscl(p1[0], st[0]), vscl(p1[1], cu));
        else st[0] = vsub(vscl(p1[0], st[0]), vscl(p1[1], cu));
        F = vadd(vscl(p2[2], F),
                 vdiv(vscl(p2[3], vsub(vscl(p0[4], st[0]), vscl(p0[5], dstore))), r->eps_mu));
        break;
      }
      case GMES_CLS_DRUDE: {              /* src/pw_drude.hh:90-101,145-150 */
        const int P = r->n1;
        const double *a = cf, *cc = cf + 3 * P;
        V *q_now = st, *q_new = st + P;
        V e_now = F;
        for (int p = 0; p < P; ++p) {
          V q_old = q_now[p];
          q_now[p] = q_new[p];
          q_new[p] = vadd(vadd(vscl(a[3 * p], q_old), vscl(a[3 * p + 1], q_now[p])),
                          vscl(a[3 * p + 2], e_now));
        }
        V sum = {0.0, 0.0};
        for (int p = 0; p < P; ++p) sum = vadd(sum, vsub(q_new[p], q_now[p]));
        V cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
        F = vadd(vadd(vscl(cc[0], cu), vscl(cc[1], sum)), vscl(cc[2], e_now));
        break;
      }
      case GMES_CLS_DCP_ADE: {            /* src/pw_dcp.hh:105-167,200-225 */
        const int D = r->n1, C = r->n2;
        const double *a = cf, *b = cf + 3 * D, *cc = cf + 3 * D + 5 * C;
        V *e_old = st, *q_old = st + 1, *q_now = st + 1 + D;
        V *p_old = st + 1 + 2 * D, *p_now = st + 1 + 2 * D + C;
        V e_now = F;
        V ds = {0.0, 0.0}, cs = {0.0, 0.0};
        for (int p = 0; p < D; ++p)
          ds = vadd(ds, vsub(vscl(1 - a[3 * p + 1], q_now[p]), vscl(a[3 * p], q_old[p])));
        for (int p = 0; p < C; ++p)
          cs = vadd(cs, vsub(vscl(1 - b[5 * p + 1], p_now[p]), vscl(b[5 * p], p_old[p])));
        V cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
        V e_new = vadd(vadd(vadd(vscl(cc[0], cu), vscl(cc[1], vadd(ds, cs))), vscl(cc[2], *e_old)),
                       vscl(cc[3], e_now));
        V mix = vadd(vadd(*e_old, vscl(2.0, e_now)), e_new);
        for (int p = 0; p < D; ++p) {
          V qn = vadd(vadd(vscl(a[3 * p], q_old[p]), vscl(a[3 * p + 1], q_now[p])),
                      vscl(a[3 * p + 2], mix));
          q_old[p] = q_now[p];
          q_now[p] = qn;
        }
        for (int p = 0; p < C; ++p) {
          V pn = vadd(vadd(vadd(vadd(vscl(b[5 * p], p_old[p]), vscl(b[5 * p + 1], p_now[p])),
                                vscl(b[5 * p + 2], *e_old)),
                           vscl(b[5 * p + 3], e_now)),
                      vscl(b[5 * p + 4], e_new));
          p_old[p] = p_now[p];
          p_now[p] = pn;
        }
        *e_old = e_now;
        F = e_new;
        break;
      }
      case GMES_CLS_DCP_PLRC: {           /* src/pw_dcp.hh:467-518,564-586 */
        const int D = r->n1, C = r->n2;
        const double *a = cf, *b = cf + 3 * D, *cc = cf + 3 * D + 6 * C;
        V *pdp = st, *pcr = st + D, *pci = st + D + C;
        V e_now = F;
        double sdr = 0.0, sdi = 0.0;
        for (int p = 0; p < D; ++p) sdr = sdr + pdp[p].re;
        for (int p = 0; p < D; ++p) sdi = sdi + pdp[p].im;
        V scr = {0.0, 0.0}, sci = {0.0, 0.0};
        for (int p = 0; p < C; ++p) scr = vadd(scr, pcr[p]);
        for (int p = 0; p < C; ++p) sci = vadd(sci, pci[p]);
        V psi = {sdr + scr.re, sdi + sci.re};   /* .real() of the imaginary sum: quirk */
        V cu = vsub(vdiv(D1, d1), vdiv(D2, d2));
        V e_new;
        if (s->lanes == 2) {
          e_new = vadd(vadd(vscl(cc[0], cu), vscl(cc[1], e_now)), vscl(cc[2], psi));
        } else {
          e_new.re = (cc[0] * cu.re + cc[1] * e_now.re) + cc[2] * psi.re;
          e_new.im = cc[1] * e_now.im + cc[2] * psi.im;
        }
        for (int p = 0; p < D; ++p) {
          pdp[p].re = (a[3 * p] * e_new.re + a[3 * p + 1] * e_now.re) + a[3 * p + 2] * pdp[p].re;
          pdp[p].im = (a[3 * p] * e_new.im + a[3 * p + 1] * e_now.im) + a[3 * p + 2] * pdp[p].im;
        }
        for (int p = 0; p < C; ++p) {
          V b0 = {b[6 * p], b[6 * p + 1]}, b1 = {b[6 * p + 2], b[6 * p + 3]};
          V b2 = {b[6 * p + 4], b[6 * p + 5]};
          pcr[p] = vadd(vadd(vscl(e_new.re, b0), vscl(e_now.re, b1)), cmul(b2, pcr[p]));
          pci[p] = vadd(vadd(vscl(e_new.im, b0), vscl(e_now.im, b1)), cmul(b2, pci[p]));
        }
        F = e_new;
        if (s->lanes == 1) F.im = 0.0;
        break;
      }
      case GMES_CLS_DM2: {                /* src/pw_dm2.hh:345-407 (real only) */
        const int N = r->n1;
        const double *omega = cf, *n_atom = cf + N;
        const double rho30 = cf[2 * N], gamma = cf[2 * N + 1], t1 = cf[2 * N + 2];
        const double t2 = cf[2 * N + 3], hbar = cf[2 * N + 4], rtol = cf[2 * N + 5];
        const double t = (n + 0.5) * dt;
        double *u = (double *)malloc(sizeof(double) * 3 * N * 4);
        double *u_new = u + 3 * N, *u_tmp = u + 6 * N, *scr = u + 9 * N;
        double av[64], bv[64];
        for (int p = 0; p < N && p < 64; ++p) {
          av[p] = n_atom[p] * gamma / t2 * exp(-t / t2);
          bv[p] = n_atom[p] * gamma * omega[p] * exp(-t / t2);
        }
        const double c_plus = 2 / hbar * gamma * exp(-t * (1 / t1 - 1 / t2));
        const double c_minus = 2 / hbar * gamma * exp(-t * (1 / t2 - 1 / t1));
        const double dd = 2 / hbar * gamma * rho30 * exp(t / t2);
        for (int p = 0; p < 3 * N; ++p) { u[p] = st[p].re; u_new[p] = u[p]; }
        double e_new = F.re;
        const double e_old = F.re;
        const double h_d = D2.re / d2;
        double error;
        int iteration = 0, fail = 0;
        do {
          double e_tmp = e_new;
          for (int p = 0; p < 3 * N; ++p) u_tmp[p] = u_new[p];
          e_new = e_old - dt * h_d;
          for (int p = 0; p < N; ++p) {
            e_new -= .5 * dt * av[p] * (u_new[3 * p] + u[3 * p]);
            e_new += .5 * dt * bv[p] * (u_new[3 * p + 1] + u[3 * p + 1]);
          }
          for (int p = 0; p < N; ++p) {
            u_new[3 * p] = u[3 * p] + .5 * dt * omega[p] * (u_new[3 * p + 1] + u[3 * p + 1]);
            u_new[3 * p + 1] = u[3 * p + 1] - .5 * dt * omega[p] * (u_new[3 * p] + u[3 * p]) +
                               .25 * dt * c_plus * (u_new[3 * p + 2] + u[3 * p + 2]) * (e_new + e_old) +
                               .5 * dt * dd * (e_new + e_old);
            u_new[3 * p + 2] = u[3 * p + 2] -
                               .25 * dt * c_minus * (u_new[3 * p + 1] + u[3 * p + 1]) * (e_new + e_old);
          }
          error = dm2_rel_error(e_new, u_new, e_tmp, u_tmp, N, scr);
          if (isnan(error)) { fail = 1; break; }
          if (++iteration >= 100 && error > rtol) { fail = 2; break; }
        } while (error > rtol);
        if (fail) {
          free(u);
          s->err = ocore_set_error(GMES_ERUNTIME, fail == 1 ? "Dm2 corrector produced an invalid error"
                                                            : "Dm2 corrector failed to converge");
          return s->err;
        }
        F.re = e_new; F.im = 0.0;
        for (int p = 0; p < 3 * N; ++p) { st[p].re = u_new[p]; st[p].im = 0.0; }
        free(u);
        break;
      }
      default:
        break;
      }
      ocore_set_flat(s, c, x, F);
    }
  }
  return GMES_OK;
}
