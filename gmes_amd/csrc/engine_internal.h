// Internal types of the MI355X FDTD engine (see include/gmes_engine.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gmes_engine.h"

namespace gmes_engine {

// ---------------------------------------------------------------------------
// scalar types: real R or interleaved complex (same layout as numpy complex)
template <class R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

template <class T> struct real_of { using type = T; };
template <class R> struct real_of<cplx<R>> { using type = R; };
template <class T> struct is_cplx { static constexpr bool value = false; };
template <class R> struct is_cplx<cplx<R>> { static constexpr bool value = true; };

// Device-side record: the FP64 host record reduced to what a kernel reads.
template <class R>
struct alignas(16) DevRec {
  int32_t cls, n1, n2, prof;
  int64_t coef_off;  // into the device coefficient pool (R)
  R c0;              // dt / eps_mu
  R eps_mu;
  R inv_eps;         // 1 / eps_mu
  R val_re, val_im;  // CONST value
};

// One grouped cell of the source table: ops [begin, end) act on (comp, cell).
struct SrcGroup {
  int32_t comp;
  int32_t begin, end;
  int32_t pad_;
  int64_t cell;
};

template <class R>
struct DevSrcOp {
  int32_t kind, wave, aux, aux_comp;
  int64_t i0, i1;
  R k, r0, r1;
};

// Compact list of the cells of one ADE class in one component.
struct AdeList {
  int64_t n = 0;
  int32_t nstate = 0;
  int32_t* cell = nullptr;   // flat index (int32: components < 2^31 cells)
  uint16_t* rec = nullptr;
  void* state = nullptr;     // SoA [nstate][n] of T
};

struct Timing {
  bool on = false;
  double ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t launches[6] = {0, 0, 0, 0, 0, 0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> pool;
};

}  // namespace gmes_engine

// The sim object behind the opaque handle.
struct gmes_sim {
  gmes_grid_desc g;
  int device = 0;
  int esize = 8;             // bytes per field element
  bool complex_ = false;
  bool fp64 = true;
  int64_t shape[6][3];
  int64_t ncell[6];
  int32_t pitch[6];          // device row pitch (elements)
  bool packed = false;       // rows unpadded: 3-D sims with rows under 256 cells
  int32_t koff[6];           // device k offset (see FieldArgs)
  int64_t dsize[6];          // device elements per component
  void* f[6] = {};           // device fields
  uint16_t* ids[6] = {};     // device ids
  uint8_t* ids8[6] = {};     // byte copies for the 3-D main kernels (at most 256 records)
  int32_t prun[3] = {};      // per axis: run (lo | hi << 16) where every CPML profile's c is zero
  bool pp_ok[2] = {false, false};  // E, H: the 3-D FP32 kernels may prefetch the set-0 CPML rows
  bool ur_ok[2] = {false, false};  // E, H: the same for the set-0 UPML rows
  int64_t prof_off0[6][3] = {};  // pool offsets of profile set 0 (component, axis)
  uint32_t upml_pre = 0;     // components whose UPML state the 3-D kernels prefetch (bit c)
  int32_t ubox[3] = {};      // per axis (lo | hi << 16): a box holding no UPML cell of those
  std::vector<uint16_t> host_ids[6];
  std::vector<gmes_mat_rec> recs;
  std::vector<double> coef;
  void* d_recs = nullptr;    // DevRec<R>[]
  void* d_coef = nullptr;    // R[]
  // profiles: host copies, then one device buffer of R with offsets
  std::vector<double> prof_host[64][6][3];
  int32_t prof_w[64][6][3] = {};
  int32_t nprof = 0;
  void* d_prof = nullptr;
  int64_t* d_prof_off = nullptr;  // [64*18]
  // full-grid state of CPML (2 words) and UPML (1 word) per component
  void* pml_state[6] = {};
  bool has_cls[6][GMES_NCLS] = {};
  gmes_engine::AdeList ade[6][GMES_NCLS];
  bool committed = false;
  // sources
  std::vector<gmes_src_op> ops;
  int32_t nwaves = 0;
  void* d_ops = nullptr;     // DevSrcOp<R>[]
  gmes_engine::SrcGroup* d_groups = nullptr;
  int32_t ngroups_e = 0, ngroups_h = 0;  // E groups first, then H groups
  double* d_waves = nullptr;
  int64_t wave_cap = 0, wave_step0 = 0, wave_n = 0;
  // aux sims
  gmes_sim* aux[8] = {};
  int32_t naux = 0;
  gmes_sim* parent = nullptr;  // the sim whose stream advances this aux sim
  hipStream_t batch_stream = nullptr;  // stream of the batch this sim belongs to
  // E ghost planes written by the 3-D E kernel (no standalone E ghost copy);
  // ade_on_ghost_source: a dispersive E cell lies on plane 0 of a
  // transverse axis (set by commit), which rules that out
  bool fuse_e = false, ade_on_ghost_source = false, allow_fuse = true;
  // fuse_hz: the 2-D (TMz) E kernel writes the z ghost plane of Hx and Hy
  bool fuse_hz = false;
  bool aux_lost = false;  // an attached auxiliary grid was destroyed
  uint64_t gen = 0;            // bumped whenever captured launch arguments go stale
  void* d_aux_ptrs = nullptr;  // [8][6] device pointers of aux fields
  // probes
  struct Probe { int32_t comp; int64_t cell; };
  std::vector<Probe> probes;
  int64_t* d_probe_cells = nullptr;
  int32_t* d_probe_comps = nullptr;
  double* d_probe_buf = nullptr;   // [nprobe][cap][2]
  int64_t probe_cap = 0, probe_base = 0, probe_n = 0;  // recorded steps since base
  std::vector<int64_t> probe_read;  // per probe: values read since probe_base
  // running-DFT monitors: boxes of storage indices, FP64 accumulators
  struct Dft {
    int32_t comp;
    int64_t lo[3], n[3];
    int32_t nfreq, decim;
    int64_t ncells;
    double* acc;
  };
  std::vector<Dft> dfts;
  double* d_dft_w = nullptr;       // [row][nfreq][2][2]
  int64_t dft_w_cap = 0, dft_step0 = 0, dft_n = 0;
  int32_t dft_nfreq = 0;
  void* d_dft_jobs = nullptr;      // DftJob<T>[dft_jobs_cap]
  int64_t dft_jobs_cap = 0;
  int64_t dft_blocks = 0;
  // time
  int64_t step = 0;          // steps enqueued
  int64_t* d_clock = nullptr;
  int32_t* d_err = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  // graph
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_dirty = true;
  // launch shape
  int tile_bx = 0, tile_by = 0, tile_ic = 0;
  gmes_engine::Timing timing;
  int64_t bytes = 0;
  // partial sums of gmes_sim_field_sum_sq (grown on demand)
  double* d_sq = nullptr;
  int64_t sq_cap = 0;
  // contiguous staging of one component for reads and writes of padded rows
  void* d_stage = nullptr;
  int64_t stage_cap = 0;  // bytes
};
