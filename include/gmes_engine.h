/*
 * gmes_engine.h -- C-ABI of the MI355X FDTD field-update engine.
 *
 * The unit behind this boundary is one simulation ("sim"): the six Yee
 * field components on the staggered storage shapes of the reference
 * (gmes/geometry.py:276-358), a per-component material-ID grid, an FP64
 * material record table, per-axis absorbing-layer profiles, and an
 * ordered table of source operations.  gmes_sim_step() advances the sim
 * by whole leapfrog steps in the order of gmes/fdtd.py:1099-1122.
 *
 * Reference interfaces replaced (one C-ABI call stands for a family of
 * SWIG/Python calls of the reference):
 *   PwMaterial<T>::attach            src/pw_material.hh:62-64, and
 *   PwMaterial<T>::merge             src/pw_material.hh:82-83
 *       -> gmes_sim_set_materials + gmes_sim_set_material_ids + gmes_sim_set_profile
 *   PwMaterial<T>::update_all        src/pw_material.hh:66-73 (called from
 *       FDTD.update_<c>, gmes/fdtd.py:673-809) -> gmes_sim_step
 *   FDTD.talk_with_<c>_neighbors     gmes/fdtd.py:811-1097 -> folded into gmes_sim_step
 *   PwSource.attach/update_all       gmes/pw_source.py:18-46 -> gmes_sim_set_sources,
 *       gmes_sim_set_waveforms
 *   TotalFieldScatteredField.step    gmes/source.py:424-425 -> gmes_sim_attach_aux
 *   Probe.write                      gmes/file_io.py:47-48 -> gmes_sim_add_probe/read_probe
 *   numpy arrays passed INPLACE_ARRAY3 (src/pw_material.i:53-84)
 *       -> gmes_sim_write_field / gmes_sim_read_field
 *
 * Conventions: every function returns GMES_OK (0) or an error code; the
 * message of the last error of the calling thread is gmes_last_error().
 * Error codes mirror the reference's exception mapping
 * (src/pw_material.i:23-33): EINVAL->ValueError, ERANGE->IndexError,
 * ETYPE->TypeError, ERUNTIME->RuntimeError.  Calls on distinct sims may
 * run concurrently from different threads; one sim is not re-entrant.
 */
#ifndef GMES_ENGINE_H_
#define GMES_ENGINE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMES_ABI_VERSION 1

/* error codes */
#define GMES_OK        0
#define GMES_EINVAL    1
#define GMES_ERANGE    2
#define GMES_ETYPE     3
#define GMES_ERUNTIME  4
#define GMES_ENOMEM    5
#define GMES_EDEVICE   6

/* storage types */
#define GMES_F32   0
#define GMES_F64   1
#define GMES_C64   2
#define GMES_C128  3

/* components, in the reference's order Ex, Ey, Ez, Hx, Hy, Hz */
#define GMES_EX 0
#define GMES_EY 1
#define GMES_EZ 2
#define GMES_HX 3
#define GMES_HY 4
#define GMES_HZ 5

/* material classes (the per-cell update rule) */
#define GMES_CLS_DUMMY       0  /* src/pw_dummy.hh:62-72: no update            */
#define GMES_CLS_CONST       1  /* src/pw_const.hh:96,198: F = value           */
#define GMES_CLS_DIELECTRIC  2  /* src/pw_dielectric.hh:114-119,288-293        */
#define GMES_CLS_CPML        3  /* src/pw_cpml.hh:127-132,339-344              */
#define GMES_CLS_UPML        4  /* src/pw_upml.hh:124-128,328-332              */
#define GMES_CLS_DRUDE       5  /* src/pw_drude.hh:90-101,145-150; also Lorentz
                                   src/pw_lorentz.hh:96-107,150-156           */
#define GMES_CLS_DCP_ADE     6  /* src/pw_dcp.hh:105-167,200-225               */
#define GMES_CLS_DCP_PLRC    7  /* src/pw_dcp.hh:467-518,564-586 (also DcpRc)  */
#define GMES_CLS_DM2         8  /* src/pw_dm2.hh:345-407                       */
#define GMES_NCLS            9

/* source operation kinds (applied after the materials of a component,
 * gmes/pw_source.py:74-87,106-119,216-486) */
#define GMES_SRC_SET      0   /* F = w                       (hard E/H source)  */
#define GMES_SRC_SUB      1   /* F = F - w                   (J/M current)      */
#define GMES_SRC_AUX_ADD  2   /* F = F + k*(r0*a[i0]+r1*a[i1]) (TFSF face)      */
#define GMES_SRC_AUX_SUB  3   /* F = F - k*(r0*a[i0]+r1*a[i1]) (TFSF face)      */

typedef struct gmes_sim gmes_sim;
typedef struct gmes_batch gmes_batch;

/* Grid description.  nx, ny, nz is the reference's whole_field_size
 * (gmes/geometry.py:145-150); component shapes follow from it and from the
 * active mask: Ex (nx, ny+1, nz+1), Ey (nx+1, ny, nz+1), Ez (nx+1, ny+1, nz),
 * Hx (nx, ny+1, nz+1), Hy (nx+1, ny, nz+1), Hz (nx+1, ny+1, nz); an inactive
 * component is (1,1,1) and reads as zero.
 * phase[c][a] is the factor of the ghost copy of component c along axis a
 * (gmes/fdtd.py:822-836): 1 for real fields, exp(+-i k_a L_a) for Bloch. */
typedef struct {
  int32_t nx, ny, nz;
  int32_t dtype;            /* GMES_F32 .. GMES_C128                 */
  uint32_t active;          /* bit c set: component c is updated     */
  int32_t device;           /* HIP device ordinal                    */
  double dx, dy, dz, dt;
  double phase[6][3][2];    /* (re, im)                              */
} gmes_grid_desc;

/* One material record.  Coefficients are FP64, computed on the host by the
 * reference's setup formulas (src/material.pyx).  coef_off indexes the
 * coefficient pool passed with the table; its layout per class:
 *   DRUDE   : a[n1][3], c[3]; n2 = 1 marks a Lorentz medium (same update)
 *   DCP_ADE : a[n1][3], b[n2][5], c[4]
 *   DCP_PLRC: a[n1][3], b[n2][3] as (re,im) pairs, c[3]
 *   DM2     : omega[n1], n_atom[n1], rho30, gamma, t1, t2, hbar, rtol
 * prof selects a profile set (gmes_sim_set_profile) for CPML/UPML.  */
typedef struct {
  int32_t cls;
  int32_t n1, n2;
  int32_t prof;
  int64_t coef_off;
  double eps_mu;            /* eps_inf (E) or mu_inf (H) of the update */
  double value[2];          /* CONST value (re, im)                    */
} gmes_mat_rec;

/* One source operation on one cell.  Operations are applied in table
 * order for a given cell; distinct cells are independent. */
typedef struct {
  int32_t comp;
  int32_t kind;             /* GMES_SRC_*                              */
  int64_t cell;             /* flat C-order index into component comp  */
  int32_t wave;             /* waveform row (SET/SUB); for AUX_* -1, or
                               the row of a real envelope that scales
                               each aux sample (Gaussian beam turn-on)  */
  int32_t aux;              /* aux slot (AUX_*)                        */
  int32_t aux_comp;         /* component of the aux sim read (AUX_*)   */
  int32_t pad_;
  int64_t i0, i1;           /* flat indices into the aux component     */
  double k, r0, r1;         /* AUX_*: F +-= k * (r0*aux[i0] + r1*aux[i1]) */
} gmes_src_op;

/* ---- library ---------------------------------------------------------- */

/* Message of the last failed call on this thread. */
const char* gmes_last_error(void);
/* ABI version (GMES_ABI_VERSION). */
int gmes_abi_version(void);
/* Number of usable devices; 0 when there is none. */
int gmes_device_count(int32_t* count);

/* ---- sim lifetime ----------------------------------------------------- */

/* Allocate the fields (zeroed) on the device given by desc->device.
 * GMES_EINVAL, before any device call, when a component would hold 2^31
 * device elements or more (kernels use 32-bit offsets). */
int gmes_sim_create(const gmes_grid_desc* desc, gmes_sim** sim);
int gmes_sim_destroy(gmes_sim* sim);
/* Shape of component comp (inactive -> 1,1,1). */
int gmes_sim_shape(const gmes_sim* sim, int32_t comp, int64_t shape[3]);
/* Bytes of device memory held by the sim (fields, state, tables). */
int gmes_sim_footprint(const gmes_sim* sim, int64_t* bytes);

/* ---- materials (replace PwMaterial attach/merge) ---------------------- */

int gmes_sim_set_materials(gmes_sim* sim, const gmes_mat_rec* recs, int32_t nrec,
                           const double* coef, int64_t ncoef);
/* ids: one uint16 record index per element of component comp, C order. */
int gmes_sim_set_material_ids(gmes_sim* sim, int32_t comp, const uint16_t* ids,
                              int64_t count);
/* Profile of set prof for component comp along axis: vals[n][width], one
 * row per storage index of comp along axis.  width 3 = CPML (b, c, kappa),
 * width 6 = UPML (c1..c6). */
int gmes_sim_set_profile(gmes_sim* sim, int32_t prof, int32_t comp, int32_t axis,
                         const double* vals, int32_t n, int32_t width);
/* Finish setup: classify cells, allocate auxiliary state (zeroed).  */
int gmes_sim_commit(gmes_sim* sim);

/* ---- sources (replace PwSource) --------------------------------------- */

int gmes_sim_set_sources(gmes_sim* sim, const gmes_src_op* ops, int64_t nops,
                         int32_t nwaves);
/* Waveform samples for steps [step0, step0+nsteps): values[w][s][re,im].
 * Row w is read by ops with wave == w at absolute step step0+s. */
int gmes_sim_set_waveforms(gmes_sim* sim, int64_t step0, int64_t nsteps,
                           const double* values);
/* Attach a 1-D auxiliary sim (TFSF incident grid) advanced between the E and
 * H half steps of sim; AUX_* ops read its fields.  Returns the slot.  The
 * aux sim keeps its own step count (it may have been run ahead). */
int gmes_sim_attach_aux(gmes_sim* sim, gmes_sim* aux, int32_t* slot);

/* ---- probes ----------------------------------------------------------- */

int gmes_sim_add_probe(gmes_sim* sim, int32_t comp, int64_t cell, int32_t* id);
/* Values recorded since the previous read: out[n][re,im]. */
int gmes_sim_read_probe(gmes_sim* sim, int32_t id, double* out, int64_t max,
                        int64_t* n);

/* ---- running-DFT monitors ---------------------------------------------
 * No reference counterpart (the reference has probes only); the flux
 * monitor of the served workload is built on these.  A monitor covers one
 * component over the box [lo, hi) of storage indices and accumulates, in
 * FP64, acc[f][cell] += w[row][f][E or H] * F(cell) at the end of every
 * step that is a multiple of decim (E values are those of n = k + 1/2, H
 * values those of n = k + 1).  All monitors of a sim share nfreq. */
int gmes_sim_add_dft(gmes_sim* sim, int32_t comp, const int64_t lo[3], const int64_t hi[3],
                     int32_t nfreq, int32_t decim, int32_t* id);
/* Weights for steps [step0, step0+nsteps): w[s][f][0 = E, 1 = H][re, im]. */
int gmes_sim_set_dft_weights(gmes_sim* sim, int64_t step0, int64_t nsteps, const double* w);
/* out[f][cell][re, im], cells of the box in C order; count = nfreq*ncells*2. */
int gmes_sim_read_dft(gmes_sim* sim, int32_t id, double* out, int64_t count);

/* ---- fields ----------------------------------------------------------- */

/* Host buffers are C-order arrays of the sim's storage type. */
int gmes_sim_write_field(gmes_sim* sim, int32_t comp, const void* host, int64_t bytes);
int gmes_sim_read_field(gmes_sim* sim, int32_t comp, void* host, int64_t bytes);
/* Everything but the sim's own six fields, as one opaque blob, for
 * pre-emption and restart: the step clock, CPML/UPML and dispersive state,
 * the running-DFT accumulators, and each attached auxiliary grid (its
 * fields and its own blob).  Restore returns the clock to the snapshot's
 * step, drops probe values not yet read and clears a deferred error; the
 * caller restores the six fields with gmes_sim_write_field.  The
 * reference's floor is FDTD.__deepcopy__ (gmes/fdtd.py:327-351), which
 * keeps fields and time only. */
int gmes_sim_state_bytes(const gmes_sim* sim, int64_t* bytes);
int gmes_sim_snapshot(gmes_sim* sim, void* host, int64_t bytes);
int gmes_sim_restore(gmes_sim* sim, const void* host, int64_t bytes);

/* Sum of |F|^2 over the box [lo, hi) of storage indices of component comp,
 * accumulated in FP64 in a fixed order (the same box gives the same bits),
 * into *out; waits for the steps enqueued before it.  Partial sums go to a
 * device buffer of the sim, grown on demand.  It stands for the
 * reference's host-side checks on the NumPy arrays, such as the regression
 * value sum(|ez|^2) of tests/test_fdtd.py:75-79, without a download. */
int gmes_sim_field_sum_sq(gmes_sim* sim, int32_t comp, const int64_t lo[3], const int64_t hi[3],
                          double* out);
/* gmes_sim_field_sum_sq without any device allocation, for callers that
 * cannot allocate (a serving loop whose other requests keep the GPU busy:
 * hipMalloc and hipFree wait for them; code that must not allocate).  No
 * counterpart in the reference.  The partial sums go to one static device
 * array per GPU (512 KiB, held until the process ends); boxes of more than
 * 2^27 cells are refused with GMES_EINVAL.  Calls on one GPU share that
 * array: the library serialises them (each returns after its result is
 * copied out), and the array keeps the last call's partial sums; every
 * call writes all the partials it reads, so nothing needs clearing.  The
 * result is bit-identical to gmes_sim_field_sum_sq's. */
int gmes_sim_field_sum_sq_noalloc(gmes_sim* sim, int32_t comp, const int64_t lo[3],
                                  const int64_t hi[3], double* out);

/* ---- time stepping ---------------------------------------------------- */

/* Enqueue nsteps full leapfrog steps on the sim's stream (asynchronous). */
int gmes_sim_step(gmes_sim* sim, int64_t nsteps);
/* Wait for the sim's stream; reports a deferred device error
 * (DM2 non-convergence / NaN -> GMES_ERUNTIME). */
int gmes_sim_sync(gmes_sim* sim);
/* Number of completed steps enqueued so far. */
int gmes_sim_get_step(const gmes_sim* sim, int64_t* step);
/* 1: capture one step as a HIP graph and replay it; 0: launch directly. */
int gmes_sim_set_graph(gmes_sim* sim, int32_t enable);
/* Launch-shape override (0 = autotuned default) for the main E/H kernels:
 * a bx x by tile of one plane (bx along k, a power of two, bx * by <= 256).
 * On a 3-D sim whose rows are under 256 cells (stored unpadded) 256 x 1,
 * the default there, means flat blocks of 256 consecutive (j, k) points.
 * On a padded 3-D sim the environment variable GMES_ICHUNK, read by this
 * call, sets the i planes per block (1, 2, 4 or 8; default 1), a tuning
 * knob the measurements keep at 1. */
int gmes_sim_set_tile(gmes_sim* sim, int32_t bx, int32_t by);
/* Tile autotuner: times the E main kernel under each candidate tile
 * (fields and PML state are restored afterwards), keeps the fastest per
 * size class for the life of the process and sets it on the sim; returns
 * the chosen tile.  No reference counterpart: the reference's update loops
 * have no launch shape (src/cpp23_support.hh:147-197). */
int gmes_sim_autotune(gmes_sim* sim, int32_t* bx, int32_t* by);
/* Fused ghost planes.  enable = 1 (the default) lets the 3-D and the 2-D
 * (TMz, TEz) E kernels write the periodic / Bloch E ghost planes
 * themselves, in place of the copy that follows the E sources
 * (gmes/fdtd.py:811-953), whenever that gives
 * the same fields: no dispersive E cell on a source plane, no E source
 * operation on a source or ghost plane, no E probe on a ghost plane.  In
 * TMz it lets the 2-D E kernel write the z ghost plane of Hx and Hy (read
 * by no E update) in place of that part of the H copy at the start of the
 * step (gmes/fdtd.py:955-1097).  enable = 0 keeps the separate copies.
 * *fused (may be NULL) reports whether either fusion is in effect now.
 * No reference counterpart. */
int gmes_sim_set_ghost_fusion(gmes_sim* sim, int32_t enable, int32_t* fused);
/* Kernel timing on the sim's stream: enable accumulates HIP-event time per
 * kernel family; read returns ms and launch counts for
 * [E main, H main, E ADE, sources, ghost, aux].  */
int gmes_sim_profile(gmes_sim* sim, int32_t enable);
int gmes_sim_read_profile(gmes_sim* sim, double ms[6], int64_t launches[6]);

/* ---- batches (many sims, one launch per kernel kind per half step) ------
 * Members keep their own fields, tables and clocks; the main, source and
 * ghost kernels of all members run as one launch each (a device job table
 * maps blocks to members), and a step of the whole batch is one captured
 * HIP graph.  A member is not stepped on its own while in a batch.  No
 * reference counterpart: the reference advances one FDTD at a time
 * (gmes/fdtd.py:1099-1122). */

/* Group sims of one storage type on one device; each keeps its own fields. */
int gmes_batch_create(gmes_sim* const* sims, int32_t n, gmes_batch** batch);
int gmes_batch_destroy(gmes_batch* batch);
/* Advance every member by nsteps (all members must have waveforms). */
int gmes_batch_step(gmes_batch* batch, int64_t nsteps);
int gmes_batch_sync(gmes_batch* batch);
int gmes_batch_set_graph(gmes_batch* batch, int32_t enable);
/* Kernel timing of a batch's own stream (HIP events; steps run without the
 * graph while it is on): read returns ms and launch counts for
 * [E 3-D main, H 3-D main, E 2-D/1-D main, H 2-D/1-D main, dispersive,
 * the rest (ghosts, sources, probes, auxiliary grids, DFT)]. */
int gmes_batch_profile(gmes_batch* batch, int32_t enable);
int gmes_batch_read_profile(gmes_batch* batch, double ms[6], int64_t launches[6]);

#ifdef __cplusplus
}
#endif

#endif /* GMES_ENGINE_H_ */
