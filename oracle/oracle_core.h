/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * Shared CPU core of the two oracles behind include/gmes_engine.h:
 *   libgmes_oracle.so  = this core + oracle_update.c (a C restatement of the
 *                        reference's per-cell arithmetic, src/pw_*.hh);
 *   _ref/libgmes_ref.so = this core + ref_update.cc (the reference's own
 *                        PwMaterial<T>::update_all, compiled from
 *                        /root/reference/src/pw_*.hh).
 * The core restates the reference's orchestration: storage shapes
 * (gmes/geometry.py:276-358), the leapfrog step order
 * (gmes/fdtd.py:1099-1122), ghost copies (gmes/fdtd.py:811-1097),
 * source updates (gmes/pw_source.py:70-131,197-486) and probes
 * (gmes/file_io.py:47-48).  Only tests/, the smoke check and the CPU
 * baseline of bench.py may load these libraries.
 */
#ifndef GMES_ORACLE_CORE_H_
#define GMES_ORACLE_CORE_H_

#include <stdint.h>
#include "../include/gmes_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ocv;   /* one field value (im = 0 if real) */

typedef struct {
  int64_t n;          /* number of cells of the class in this component */
  int64_t *cell;      /* flat indices, ascending                        */
  int32_t *rec;       /* record index per cell                          */
  int32_t nstate;     /* state words (ocv) per cell                     */
  ocv *state;         /* [n][nstate]                                    */
} ocls_list;

struct gmes_sim {
  gmes_grid_desc g;
  int lanes;                     /* 1 real, 2 complex                 */
  int64_t shape[6][3];
  int64_t ncell[6];
  double *f[6];                  /* native layout: lanes doubles/cell */
  uint16_t *ids[6];
  gmes_mat_rec *recs; int32_t nrec;
  double *coef; int64_t ncoef;
  /* profiles: prof[p][comp][axis] -> rows x width */
  int32_t nprof;
  double *prof[64][6][3]; int32_t prof_n[64][6][3]; int32_t prof_w[64][6][3];
  ocls_list cls[6][GMES_NCLS];   /* per component, per class           */
  gmes_src_op *ops; int64_t nops;
  int32_t nwaves; int64_t wave_step0, wave_n; double *waves;
  gmes_sim *aux[8]; int32_t naux;
  struct { int32_t comp; int64_t cell; double *buf; int64_t n, cap; } probe[64];
  int32_t nprobe;
  /* running-DFT monitors (restatement of the engine's FP64 accumulation) */
  struct { int32_t comp; int64_t lo[3], n[3], ncells; int32_t decim; double *acc; } dft[16];
  int32_t ndft, dft_nfreq;
  int64_t dft_step0, dft_n; double *dft_w;
  int64_t step;                  /* completed steps                    */
  int committed;
  int err;                       /* deferred error (DM2)               */
  char errmsg[256];
  void *impl;                    /* per-backend material data          */
};

/* component geometry: in1, in2, d1 axis, d2 axis, own axis */
extern const int GMES_IN1[6], GMES_IN2[6], GMES_AX1[6], GMES_AX2[6], GMES_AX0[6];

int ocore_set_error(int code, const char *fmt, ...);
double ocore_spacing(const gmes_sim *s, int axis);
ocv ocore_get(const gmes_sim *s, int comp, int64_t i, int64_t j, int64_t k);
ocv ocore_get_flat(const gmes_sim *s, int comp, int64_t flat);
void ocore_set_flat(gmes_sim *s, int comp, int64_t flat, ocv v);

/* backend hooks (implemented by oracle_update.c or ref_update.cc) */
int obackend_commit(gmes_sim *s);            /* after ids/records/profiles */
int obackend_update(gmes_sim *s, int comp, double n);   /* all materials of comp */
void obackend_free(gmes_sim *s);

#ifdef __cplusplus
}
#endif
#endif
