// Kernel argument blocks and launchers shared by kernels.hip and engine.hip.
#pragma once

#include "engine_internal.h"

namespace gmes_engine {

constexpr int kDm2Max = 8;    // transitions per DM2 cell held in registers
constexpr int kRecLds = 128;  // records staged in LDS by the 3-D main kernel

// Per-sim control block in device memory: everything a step reads that
// changes between steps, so one captured step graph can be replayed.
struct Ctl {
  int64_t clock;       // steps started (thread 0 of the first kernel of a step adds 1)
  int64_t wave_step0;  // absolute step of waveform row 0
  int64_t wave_n;      // rows per waveform
  int64_t probe_base;  // absolute step of probe slot 0
  int64_t dft_step0;   // absolute step of DFT weight row 0
  int64_t dft_n;       // rows of DFT weights
  int32_t err;         // deferred device error (1 NaN, 2 no convergence)
  int32_t pad_;
};

constexpr int kPad = 32;  // elements before row 0 of every device array

// Device layout of a component: logical (i,j,k) of a component with shape
// (s0,s1,s2) lives at element kPad + (i*s1 + j)*p + k - koff.  The row
// pitch p is s2 rounded up to 32 elements when s2 >= 32 (else s2), and
// koff = 1 for Hx and Hy on a padded layout, so that the first updated
// cell of every row (k = 0 for E and Hz, k = 1 for Hx, Hy) starts a
// 128-byte line.  3-D sims with rows under 256 cells keep p = s2 and
// koff = 0 (gmes_sim::packed).  Material ids and CPML/UPML state use the
// same layout.
template <class T>
struct FieldArgs {
  T* f[6];
  int s[6][3];
  int p[6];
  int koff[6];
  unsigned active;
};

template <class T>
__host__ __device__ __forceinline__ int didx(const FieldArgs<T>& a, int c, int i, int j, int k) {
  return kPad + (i * a.s[c][1] + j) * a.p[c] + k - a.koff[c];
}

template <class T>
struct MainArgs {
  using R = typename real_of<T>::type;
  FieldArgs<T> fa;
  const uint16_t* ids[6];
  const DevRec<R>* recs;
  const R* prof;
  const int64_t* prof_off;
  void* pml[6];
  int dsize[6];  // device elements per component (offset of the 2nd CPML word)
  R d[3];
  R inv_d[3];
  int nrec;
  int n[3];  // nx, ny, nz
  // 1: the 3-D E kernel writes the E ghost planes itself (fused_ghost)
  int fuse;
  int gax[3][2];      // transverse axes a1, a2 of Ex, Ey, Ez
  cplx<R> gph[3][2];  // ghost phase along a1, a2
  const uint8_t* ids8[6];  // the ids as bytes (real-FP32 3-D sims), when narrow
  int narrow;
  // generic kernels: 1 when every row has at most 2 elements (kext); each
  // thread then covers a whole row (bx = 1)
  int kloop, kext;
  // real-FP32 3-D kernels: components with CPML state (bit c), and per axis
  // the coordinate run (lo | hi << 16) where every CPML profile's c is zero
  unsigned cpml_mask;
  int prun[3];
  int po0[6][3];  // offsets of profile set 0's rows (component, axis) in the pool
  // real-FP32 3-D kernels with CPML state in all three components of the
  // field type: pp = 1 when, per axis, the two components that read a
  // profile along it have identical set-0 rows; pprow[a] is then the pool
  // offset of one of them, and the kernels load the row of each axis with
  // the cell's other operands (lanes outside the zero-c run of that axis)
  int pp;
  int pprow[3];
  // real-FP32 3-D kernels: components whose UPML state word is loaded with
  // the other operands (bit c), and a box (per axis lo | hi << 16) that
  // holds no UPML cell of those components
  unsigned upml_mask;
  int ubox[3];
  // real-FP32 3-D kernels whose three components of the field type all
  // carry prefetched UPML state: ur = 1 when, per axis, the two components
  // that read it as a curl axis have identical set-0 rows; urs[a] is then
  // the pool offset of that row and uro[a] that of the owning component's
  // row, and the kernels load both with the cell's other operands (lanes
  // outside ubox)
  int ur;
  int urs[3], uro[3];
  // 2-D modes with a collapsed z axis (TMz): 1 when the E kernel writes the
  // z ghost plane (k = 0) of Hx and Hy, which no E update reads, in place
  // of that part of the H ghost copy; hzph: the copy's phase along z for
  // Hx and Hy; hyxph: Hy's phase along x (its corner line)
  int hzg;
  cplx<R> hzph[2];
  cplx<R> hyxph;
};

template <class T>
struct AdeArgs {
  using R = typename real_of<T>::type;
  FieldArgs<T> fa;
  const int32_t* cell;
  const uint16_t* rec;
  void* state;
  int n;
  const DevRec<R>* recs;
  const R* coef;
  R d[3];
  R dt;
  Ctl* ctl;
};

template <class T>
struct SrcArgs {
  using R = typename real_of<T>::type;
  FieldArgs<T> fa;
  const SrcGroup* groups;
  int ngroups;
  const DevSrcOp<R>* ops;
  const double* waves;
  void* const* aux;
  Ctl* ctl;
};

template <class T>
struct ProbeArgs {
  FieldArgs<T> fa;
  const int64_t* cells;
  const int32_t* comps;
  int n;
  int electric;
  double* buf;
  int64_t cap;
  Ctl* ctl;
};

template <class T>
struct GhostArgs {
  using R = typename real_of<T>::type;
  FieldArgs<T> fa;
  int ncomp;
  int comp[3];
  int ax[3][3];      // a1, a2, own axis
  int64_t count[3];  // ghost points per component
  cplx<R> ph[3][2];  // phase of the copy along a1, a2
  int tick;
  Ctl* ctl;
  int64_t skip[3];   // leading points of a component's copy left out (written elsewhere)
};

// Batched-launch job tables: one entry per member sim; block0 is the
// member's first block in the launch (entries sorted by block0).
// 3-D and collapsed-z members run flat blocks of 256 consecutive points,
// so rows that are not a multiple of a tile width leave no lanes idle:
// 3-D, gx blocks per i plane over its row-major (j, k) points; collapsed
// z, gx blocks over the row-major (i, j) plane whose rows are bx long.
// Other members run bx x (256/bx) tiles, gx along k and gy along j per
// plane.
template <class T>
struct MainJob {
  MainArgs<T> m;
  int bx;
  int gx, gy;
  int pad_;
  int64_t block0;
};

template <class T>
struct SrcJob {
  SrcArgs<T> a;
  int64_t block0;
};

template <class T>
struct GhostJob {
  GhostArgs<T> g;
  int64_t block0;
};

// Running-DFT monitor of one component over a box of storage indices:
// acc[f][cell] += w[row][f][E or H] * F(cell) in FP64, on steps that are
// multiples of decim.
template <class T>
struct DftJob {
  T* f;
  int s1, p, koff, comp;
  int lo[3], n[3];
  int nfreq, decim, electric, pad_;
  int64_t ncells;
  double* acc;      // [nfreq][ncells][re, im]
  const double* w;  // [row][nfreq][2: E, H][re, im]
  Ctl* ctl;
  int64_t block0;
};

template <class T>
struct AdeJob {
  AdeArgs<T> a;
  int comp;
  int pad_;
  int64_t block0;
};

// Sum of |F|^2 over a box of storage indices of one component, in FP64
// and in a fixed order: pass 1 reduces kSqBlockCells cells per block into
// partial[block]; pass 2 (one block) reduces the partials into
// partial[nblocks].  The same box gives the same bits whatever buffer
// holds the partials.
constexpr int kSqBlockCells = 2048;  // 256 threads x 8 cells
template <class T>
struct SqBox {
  const T* f;    // element 0 of the device array
  int s1, p, koff;
  int lo[3], n[3];
  int ncells;    // < 2^31 (a component holds < 2^31 elements)
};
inline int64_t sq_blocks(int64_t ncells) { return (ncells + kSqBlockCells - 1) / kSqBlockCells; }
template <class T>
void launch_sum_sq(const SqBox<T>& b, double* partial, hipStream_t st);

// rows x n2 elements of esize bytes from rows spaced spitch elements apart
// to rows spaced dpitch apart (field reads and writes of padded layouts)
void launch_pitch_copy(const void* src, void* dst, int64_t rows, int n2, int spitch, int dpitch, int esize,
                       hipStream_t st);
template <class T>
void launch_main(bool electric, const MainArgs<T>& m, int ext[3], int bx, int by, hipStream_t st);
// flat blocks over packed rows (3-D; see k_main3d_flat)
template <class T>
void launch_main3d_flat(bool electric, const MainArgs<T>& m, hipStream_t st);
// collapsed-z 2-D kernels: i rows per thread (1 measured fastest of 1, 2
// and 4 at 4096^2; 1.4x the generic kernel in FP32)
constexpr int k2zRows = 1;
// kinds of batched main launch
// kBatch3dLds / kBatch2zLds: 3-D / collapsed-z jobs whose record tables
// all fit the LDS staging (RecStage::use_lds for 256-thread blocks)
// kBatch3dPP: real-FP32 3-D jobs whose records fit the LDS staging and
// whose CPML profile rows the kernel loads with the other operands
// (MainArgs::pp); kBatch3dUR: the same for UPML rows (MainArgs::ur)
constexpr int kBatchGeneric = 0, kBatch3d = 1, kBatch2z = 2, kBatch3dLds = 3, kBatch2zLds = 4, kBatch3dPP = 5,
              kBatch3dUR = 6;
template <class T>
void launch_main_batch(bool electric, int kind, const MainJob<T>* jobs, int njobs, int64_t nblocks,
                       hipStream_t st);
template <class T>
void launch_src_batch(const SrcJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st);
template <class T>
void launch_ade_batch(int cls, const AdeJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st);
template <class T>
void launch_dft_batch(const DftJob<T>* jobs, int njobs, int64_t nblocks, hipStream_t st);
template <class T>
void launch_ghost_batch(const GhostJob<T>* jobs, int njobs, int64_t nblocks, int tick, hipStream_t st);
template <class T>
void launch_main2z(bool electric, const MainArgs<T>& m, int ext[3], hipStream_t st);
template <class T>
void launch_main3d(bool electric, const MainArgs<T>& m, int bx, int by, int ichunk, hipStream_t st);
template <class T>
void launch_ade(int cls, int c, const AdeArgs<T>& m, hipStream_t st);
template <class T>
void launch_src(const SrcArgs<T>& m, hipStream_t st);
template <class T>
void launch_probe(const ProbeArgs<T>& m, hipStream_t st);
template <class T>
void launch_ghost(const GhostArgs<T>& m, int64_t total, hipStream_t st);

}  // namespace gmes_engine
