// Host-side declarations of the kernel launchers (implemented in the .hip files).
#pragma once
#include <hip/hip_runtime.h>

#include "sbp_device.hpp"

namespace sbp {

// Configurations of the fused stage kernel: TX points along ξ, TY along η,
// each thread owning RX adjacent ξ-columns by RY consecutive η-rows, and the
// variant of keeping the dissipation terms between passes.
struct TileShape {
  int tx, ty, rx, ry, var;
};
constexpr int kNumTileShapes = 13;
extern const TileShape kTileShapes[kNumTileShapes];

// Operator application along one axis (0 = ξ, 1 = η) of a (ny, nx) array.
hipError_t launch_apply_op(int op, int kind, int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny,
                           int nx, hipStream_t stream);

// Fused stage kernel over `ntiles` tiles of blocks sharing one operator pair.
// mode: StageMode; edge: the tiles lie on block edges and add the edge SATs.
// Returns hipErrorInvalidValue for an unsupported combination and
// hipErrorNotSupported for a configuration not built for the pair.
hipError_t launch_stage(int op_eta, int op_xi, bool upwind, bool edge, int shape, int mode, const StageArgs& args,
                        int ntiles, hipStream_t stream);
// Launch bounds (threads per workgroup) and LDS bytes for a tile shape.
int stage_threads(int shape);

// Pointwise metrics from the four derivative planes (grid.rs:110-131).
hipError_t launch_metrics_finish(const double* dx_dxi, const double* dx_deta, const double* dy_dxi, const double* dy_deta,
                                 double* metrics, int64_t ps, int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream);
// Characteristic SATs of the listed edges at stage `mode` into BlockDev::sat
// (maxlen: the longest listed edge).
hipError_t launch_edge_sat(const StageArgs& a, const SatSeg* segs, int nsegs, int maxlen, int mode, hipStream_t stream);
// Mixed form: reference face terms of the `ntiles` tiles at a.tiles (tile
// shape `shape`) from the state stage `mode` reads, into a.mixref.
hipError_t launch_mix_ref(const StageArgs& a, int shape, int mode, int ntiles, hipStream_t stream);
// Expressions of `prog` (a device copy) on a block at time t into 4 planes.
hipError_t launch_eval_field(const expr::Program* prog, double t, const double* xc, const double* yc, double* out, int64_t ps,
                             int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream);
// Boundary data of the listed Eval edges at the stage time of `a` into BlockDev::edata.
hipError_t launch_eval_edges(const StageArgs& a, const SatSeg* segs, int nsegs, int maxlen, hipStream_t stream);
// Metric pairs (BlockDev::mpair, 4·ps doubles) from the 5 metric planes.
hipError_t launch_metric_pairs(const double* metrics, int64_t ps, int64_t ld, int ny, int nx, double* pairs,
                               hipStream_t stream);
// Vortex initial state on (ny, nx) coordinates into 4 planes.
hipError_t launch_vortex(const VortexDev& vp, double gamma, double t, const double* xc, const double* yc, double* out,
                         int64_t ps, int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream);
// Compensated H-norm sums (Field::h2_err, euler/src/lib.rs:377-432):
// Σ_c Σ_ij w_ij (a_c - b_c)² with w = hx_i hy_j, b = the vortex at time t on
// (xc, yc) when use_vortex, b = 0 when b is null. A grid of
// reduce_workgroups(ny) <= kReduceMaxWG workgroups strides over the rows;
// scratch holds kReduceScratch doubles: per-workgroup (sum, compensation)
// pairs, then the result pair at scratch[2·kReduceMaxWG].
constexpr int kReduceMaxWG = 1024;
constexpr int kReduceScratch = 2 * kReduceMaxWG + 2;
int reduce_workgroups(int ny);
struct H2Args {
  const double* a;
  int64_t psa, lda;
  const double* b;
  int64_t psb, ldb;
  VortexDev vp;
  int use_vortex;
  double gamma, t;
  const double* xc;
  const double* yc;
  int64_t ldc;
  int ny, nx;
};
hipError_t launch_h2_sum(int op_eta, int op_xi, const H2Args& args, double* scratch, hipStream_t stream);
// advance_adaptive: per block, the compensated sum of its tiles' (sum,
// compensation) pairs (BlockDev::tile_off, ntx·nty slots) into out[2b], out[2b+1].
hipError_t launch_block_sums(const BlockDev* blocks, int nblocks, const double* tile_sums, double* out, hipStream_t stream);
// max |û| and max |v̂| (system.rs:561-578) into scratch[2·kReduceMaxWG],
// scratch[2·kReduceMaxWG + 1] (kReduceScratch doubles of scratch).
hipError_t launch_maxspeed(const double* q, const double* metrics, int64_t ps, int64_t ld, int ny, int nx, double* scratch,
                           hipStream_t stream);
// Characteristic SAT of one edge added to k (SAT_north/south/east/west).
struct SatEdgeArgs {
  int dir;
  double* k;
  const double* y;
  const double* met;
  const double* g;
  int64_t ps, ld;
  int ny, nx;
  double hi, gamma;
};
hipError_t launch_sat_edge(const SatEdgeArgs& args, hipStream_t stream);
// time += dt on the device.
hipError_t launch_advance_time(double* time, double dt, hipStream_t stream);
// out = prev, then out = out + s_j·k_j for j < nk in order, unfused (the
// reference's clone_from + scaled_add, integrate/src/lib.rs:187-211) over the
// 4 planes of one block; out may be prev.
struct LinComb {
  const double* k[kMaxWorkSets];
  double s[kMaxWorkSets];
  int nk;
};
hipError_t launch_lincomb(double* out, const double* prev, const LinComb& lc, int64_t ps, int64_t ld, int ny, int nx,
                          hipStream_t stream);
// 2:1 interpolation of a vector (InterpolationOperator; one thread per output).
hipError_t launch_interpolate(int interp_op, bool fine2coarse, const double* in, int nin, double* out, int nout,
                              hipStream_t stream);
// Copy between pitched planes (4 or 5 planes) on the device.
hipError_t launch_copy_planes(const double* src, int64_t sps, int64_t sld, double* dst, int64_t dps, int64_t dld, int nplanes,
                              int ny, int nx, hipStream_t stream);

}  // namespace sbp
