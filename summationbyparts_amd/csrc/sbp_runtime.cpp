// Host runtime of libsbpgpu: contexts, systems (blocks, boundary conditions,
// device buffers), fused RK4 stepping captured in HIP graphs, the tile-shape
// autotuner and the C ABI declared in include/sbpgpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/sbpgpu.h"
#include "sbp_device.hpp"
#include "sbp_interp.hpp"
#include "sbp_kernels.hpp"
#include "sbp_operators.hpp"

namespace sbp {
std::shared_ptr<const std::vector<unsigned char>> expr_blob(sbp_expr e);
hipError_t launch_stage_opxi_0(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);
hipError_t launch_stage_opxi_1(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);
hipError_t launch_stage_opxi_2(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);
hipError_t launch_stage_opxi_3(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);
hipError_t launch_stage_opxi_4(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);
hipError_t launch_stage_opxi_5(int, bool, bool, int, int, const StageArgs&, int, hipStream_t);

hipError_t launch_stage(int op_eta, int op_xi, bool upwind, bool edge, int shape, int mode, const StageArgs& args,
                        int ntiles, hipStream_t stream) {
  if (ntiles <= 0) return hipSuccess;
  switch (op_xi) {
    case 0: return launch_stage_opxi_0(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    case 1: return launch_stage_opxi_1(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    case 2: return launch_stage_opxi_2(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    case 3: return launch_stage_opxi_3(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    case 4: return launch_stage_opxi_4(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    case 5: return launch_stage_opxi_5(op_eta, upwind, edge, shape, mode, args, ntiles, stream);
    default: return hipErrorInvalidValue;
  }
}

int stage_threads(int shape) {
  const TileShape& t = kTileShapes[shape];
  return (t.tx / t.rx) * (t.ty / t.ry);
}
}  // namespace sbp

using namespace sbp;

namespace {

struct OpInfo {
  int M, N, D, h2, upwind;
  const double* h;
};

template <class OP>
OpInfo info_of() {
  return {OP::M, OP::N, OP::D, OP::is_h2, OP::is_upwind, OP::h};
}

bool op_info(int op, OpInfo& out) {
  switch (op) {
    case UPWIND4: out = info_of<ops::Upwind4>(); return true;
    case UPWIND4H2: out = info_of<ops::Upwind4h2>(); return true;
    case UPWIND9: out = info_of<ops::Upwind9>(); return true;
    case UPWIND9H2: out = info_of<ops::Upwind9h2>(); return true;
    case SBP4: out = info_of<ops::SBP4>(); return true;
    case SBP8: out = info_of<ops::SBP8>(); return true;
    default: return false;
  }
}

template <class OP>
void coefs_of(int kind, double* start, double* diag, double* end, bool& ok) {
  ok = true;
  const int M = OP::M, N = OP::N, D = OP::D;
  if (kind == SBP_DIFF) {
    for (int i = 0; i < M; ++i)
      for (int k = 0; k < N; ++k) {
        start[i * N + k] = OP::diff_start[i][k];
        end[i * N + k] = -OP::diff_start[M - 1 - i][N - 1 - k];
      }
    for (int d = 0; d < D; ++d) diag[d] = OP::diff_diag[d];
  } else {
    if constexpr (OP::is_upwind) {
      for (int i = 0; i < M; ++i)
        for (int k = 0; k < N; ++k) {
          start[i * N + k] = OP::diss_start[i][k];
          end[i * N + k] = OP::diss_start[M - 1 - i][N - 1 - k];
        }
      for (int d = 0; d < D; ++d) diag[d] = OP::diss_diag[d];
    } else {
      ok = false;
    }
  }
}

std::mutex g_gamma_mu;
double g_gamma = 0.0;
bool g_gamma_set = false;

bool gamma_get(double& g) {
  std::lock_guard<std::mutex> lk(g_gamma_mu);
  g = g_gamma;
  return g_gamma_set;
}

int from_hip(hipError_t e) {
  if (e == hipSuccess) return SBP_OK;
  if (e == hipErrorOutOfMemory) return SBP_ERR_OOM;
  return SBP_ERR_HIP;
}

#define SBP_TRY_HIP(expr)                   \
  do {                                      \
    hipError_t e_ = (expr);                 \
    if (e_ != hipSuccess) return from_hip(e_); \
  } while (0)

#define SBP_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ != SBP_OK) return s_; \
  } while (0)

int status_from_errword(unsigned w) {
  if (!w) return SBP_OK;
  for (int b = 1; b < 32; ++b)
    if (w & (1u << b)) return b;
  return SBP_ERR_INVALID;
}

VortexDev to_dev(const sbp_vortex_params& p) {
  VortexDev v{};
  v.n = std::max(0, std::min(5, (int)p.nvortices));
  for (int i = 0; i < 5; ++i) {
    v.x0[i] = p.vortices[i].x0;
    v.y0[i] = p.vortices[i].y0;
    v.rstar[i] = p.vortices[i].rstar;
    v.eps[i] = p.vortices[i].eps;
  }
  v.mach = p.mach;
  return v;
}

int64_t pitch_of(int nx) { return ((int64_t)nx + 15) / 16 * 16; }

// The stage kernels address 5 planes of a block with 32-bit byte offsets:
// 5 planes of pitch x ny doubles stay below 2^32 bytes.
bool planes_fit(int nx, int ny) { return 5 * pitch_of(nx) * (int64_t)ny * 8 < ((int64_t)1 << 32); }

}  // namespace

// Device scratch of sbp_h2_err_noalloc: the per-workgroup partial sums and
// the result pair of one H-norm reduction (kReduceScratch doubles, 16 KiB per
// device, held until the process ends). In host code `scratch` is only a
// stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch_of), and only that
// address is passed to kernels and copies.
static __device__ __align__(256) double scratch[kReduceScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once.
hipError_t noalloc_scratch_of(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

struct sbp_context_s {
  int device = 0;
  double* noalloc_scratch = nullptr;  // this device's `scratch` (sbp_h2_err_noalloc)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // second stream of the context (default priority): the edge tiles of a
  // stage run on it next to the interior tiles; fork/join events order it
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  unsigned* d_err = nullptr;
  BlockDev* d_tmpblock = nullptr;
  TileDev* d_tmptiles = nullptr;
  int tmptiles_cap = 0;
  // autotuner cache: (op_eta, op_xi, upwind, total points, nblocks) -> shape
  std::map<std::tuple<int, int, int, int64_t, int>, int> tuned;
};

namespace {

struct HostBlock {
  int nx, ny, op_eta, op_xi;
  int64_t ld, ps;
  double* d_metrics = nullptr;
  double* d_mpair = nullptr;  // metric pairs of the stage kernel (BlockDev::mpair)
  double* d_x = nullptr;
  double* d_y = nullptr;
  double* d_buf[4] = {nullptr, nullptr, nullptr, nullptr};
  double* d_k = nullptr;
  double* d_rk = nullptr;  // Runge-Kutta work sets (rk_sets x 4 planes), allocated on first use
  int rk_sets = 0;
  double dt_scale = 1.0;  // the block steps with dt·dt_scale (independent solves in one batch)
  int bc_kind[4] = {SBP_BC_THIS, SBP_BC_THIS, SBP_BC_THIS, SBP_BC_THIS};
  int bc_nb[4] = {0, 0, 0, 0};
  int bc_iop[4] = {0, 0, 0, 0};   // SBP_BC_INTERPOLATE operator
  std::vector<int> bc_pieces[4];  // SBP_BC_MULTIGRID (block, start, end) triples
  sbp_vortex_params bc_vortex[4];
  std::shared_ptr<const std::vector<unsigned char>> bc_prog[4];  // SBP_BC_EVAL program blobs
};

struct Group {
  int op_eta, op_xi;
  bool upwind;
  std::vector<int> blocks;
  int shape = -1;
  // schedule: 0 split over two streams, 1 one dispatch, 2 the edge-SAT kernel
  // and then one dispatch (-1 not chosen yet)
  int sched = -1;
  // order of the tiles in a dispatch (tuned with the schedule): 0 = each
  // XCD gets a contiguous band (workgroup g runs on XCD g mod 8), 1 = plain
  // row-major order within each block
  int order = 0;
  // the tiles on block edges (they add the edge SATs), then the interior tiles
  int etile_offset = 0, entiles = 0;
  int tile_offset = 0, ntiles = 0;
  int sat_off = 0, nsat = 0, sat_maxlen = 0;  // the group's block edges for the edge-SAT kernel
  int eval_off = 0, neval = 0, eval_maxlen = 0;  // the group's Eval edges (same list, after the SAT edges)
};

}  // namespace

struct sbp_system_s {
  sbp_context ctx = nullptr;
  std::vector<HostBlock> blocks;
  std::vector<Group> groups;
  BlockDev* d_blocks = nullptr;
  TileDev* d_tiles = nullptr;
  int* d_pieces = nullptr;  // MultiGrid piece lists of all edges
  // per-edge SAT buffers (BlockDev::sat, schedule 2) and the boundary data of
  // the Eval edges at the current stage time (BlockDev::edata) in one device
  // array; the edge lists of the edge-SAT and Eval-data kernels, grouped by
  // operator group (Group::sat_off / nsat, eval_off / neval)
  double* d_edata = nullptr;
  SatSeg* d_evsegs = nullptr;
  // device copies of the Eval edges' expression programs (BlockDev::eprog)
  unsigned char* d_progs = nullptr;
  double* d_time = nullptr;      // system time
  double* d_time_rhs = nullptr;  // explicit time for sbp_system_rhs
  unsigned* d_err = nullptr;
  double* d_partials = nullptr;
  // advance_adaptive: per-tile error sums (2 doubles per tile of every
  // block, BlockDev::tile_off), their per-block totals, a pinned host copy
  int64_t total_tiles = 0;
  double* d_tile_sums = nullptr;
  int64_t tile_sums_cap = 0;
  double* d_block_sums = nullptr;
  double* h_block_sums = nullptr;
  int partials_cap = 0;
  // mixed form: the tiles' reference face terms (StageArgs::mixref, 16
  // doubles per tile slot)
  double* d_mixref = nullptr;
  int64_t mixref_cap = 0;
  int rhs_kind = SBP_RHS_AUTO;
  int precision = SBP_FP64;
  // mixed form: 1 while y_n lives in the accumulator buffer (odd steps)
  int parity = 0;
  int forced_shape = -1;
  int forced_schedule = -1;
  int forced_order = -1;
  bool dirty = true;  // device descriptors / tiles need rebuilding
  // captured steps: graph of `graph_steps` RK4 steps for step size graph_dt
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  double graph_dt = NAN;
  int graph_steps = 0;
};

namespace {

bool block_upwind(const sbp_system_s* s, const HostBlock& b, int* err) {
  OpInfo oe, ox;
  op_info(b.op_eta, oe);
  op_info(b.op_xi, ox);
  const bool both = oe.upwind && ox.upwind;
  *err = SBP_OK;
  if (s->rhs_kind == SBP_RHS_TRAD) return false;
  if (s->rhs_kind == SBP_RHS_UPWIND) {
    if (!both) *err = SBP_ERR_NOT_UPWIND;
    return true;
  }
  return both;
}

void destroy_graph(sbp_system_s* s) {
  if (s->graph) hipGraphExecDestroy(s->graph);
  if (s->graph_def) hipGraphDestroy(s->graph_def);
  s->graph = nullptr;
  s->graph_def = nullptr;
  s->graph_dt = NAN;
  s->graph_steps = 0;
}

int ntiles_along(int n, int t) { return n <= t ? 1 : (n + t - 1) / t; }


// Device descriptor for block i with the tile counts of `shape`.
BlockDev make_dev(const sbp_system_s* s, int i, int shape) {
  const HostBlock& h = s->blocks[i];
  BlockDev b{};
  b.nx = h.nx;
  b.ny = h.ny;
  b.ld = (int)h.ld;
  b.ps = h.ps;
  b.op_xi = h.op_xi;
  b.op_eta = h.op_eta;
  const TileShape& ts = kTileShapes[shape < 0 ? 0 : shape];
  b.ntx = ntiles_along(h.nx, ts.tx);
  b.nty = ntiles_along(h.ny, ts.ty);
  b.metrics = h.d_metrics;
  b.mpair = h.d_mpair;
  b.xc = h.d_x;
  b.yc = h.d_y;
  for (int r = 0; r < 4; ++r) b.buf[r] = h.d_buf[r];
  b.rhs_in = h.d_buf[BUF_YN];
  b.rhs_out = h.d_k;
  for (int r = 0; r < kMaxWorkSets; ++r) b.rk[r] = r < h.rk_sets ? h.d_rk + (int64_t)r * 4 * h.ps : nullptr;
  OpInfo ox, oe;
  op_info(h.op_xi, ox);
  op_info(h.op_eta, oe);
  b.idx_xi = 1.0 / (ox.h2 ? 1.0 / (double)(h.nx - 2) : 1.0 / (double)(h.nx - 1));
  b.idx_eta = 1.0 / (oe.h2 ? 1.0 / (double)(h.ny - 2) : 1.0 / (double)(h.ny - 1));
  b.hi_xi = (ox.h2 ? (double)(h.nx - 2) : (double)(h.nx - 1)) / ox.h[0];
  b.hi_eta = (oe.h2 ? (double)(h.ny - 2) : (double)(h.ny - 1)) / oe.h[0];
  b.dt_scale = h.dt_scale;
  int e;
  b.upwind = block_upwind(s, h, &e);
  for (int d = 0; d < 4; ++d) {
    b.edge_src[d] = h.bc_kind[d];
    b.edge_nb[d] = h.bc_nb[d];
    b.vortex[d] = to_dev(h.bc_vortex[d]);
    b.edge_iop[d] = h.bc_iop[d];
    // fine to coarse when this block has fewer points along ξ (lib.rs:818)
    b.edge_f2c[d] = h.bc_kind[d] == SBP_BC_INTERPOLATE && h.nx < s->blocks[h.bc_nb[d]].nx;
    b.edge_pieces[d] = nullptr;  // set by the upload, which owns the device copy
    b.edge_npieces[d] = (int)h.bc_pieces[d].size() / 3;
  }
  return b;
}

// Autotuner cache key part: RHS kind and precision of a group.
int tune_kind(const sbp_system_s* s, const Group& g) { return (int)g.upwind + 2 * (s->precision == SBP_MIXED); }

// Tile list of a group ordered for the 8 XCDs of an MI355X: workgroup g runs
// on XCD g mod 8, so the natural (block, row, column) order is cut into 8
// contiguous chunks and interleaved; each XCD then sweeps one compact band of
// rows (measured: 30% faster than the plain order at 4096²). Orders that
// visit 2-8 tile rows of a column before the next column, or narrow vertical
// strips per XCD, read the shared η halos closer together (the strips raise
// their L2 hit rate from 39% to 51%) but run 1-20% slower at 4096².
void xcd_order(const std::vector<TileDev>& nat, std::vector<TileDev>& out) {
  const int T = (int)nat.size();
  const int X = 8;
  std::vector<int> start((size_t)X + 1, 0);
  for (int x = 0; x < X; ++x) start[x + 1] = start[x] + T / X + (x < T % X ? 1 : 0);
  const int maxlen = T / X + (T % X ? 1 : 0);
  for (int j = 0; j < maxlen; ++j)
    for (int x = 0; x < X; ++x)
      if (start[x] + j < start[x + 1]) out.push_back(nat[start[x] + j]);
}

// The group's interior tiles and its tiles on block edges, each list in XCD
// band order (order 0) or row-major order (order 1).
void build_tiles(const sbp_system_s* s, const Group& g, int shape, std::vector<TileDev>& interior,
                 std::vector<TileDev>& edge) {
  const TileShape& ts = kTileShapes[shape];
  std::vector<TileDev> ni, ne;
  for (int bi : g.blocks) {
    const HostBlock& h = s->blocks[bi];
    const int kx = ntiles_along(h.nx, ts.tx), ky = ntiles_along(h.ny, ts.ty);
    for (int ty = 0; ty < ky; ++ty)
      for (int tx = 0; tx < kx; ++tx) {
        const TileDev t{bi, (int16_t)tx, (int16_t)ty};
        (tx == 0 || tx == kx - 1 || ty == 0 || ty == ky - 1 ? ne : ni).push_back(t);
      }
  }
  if (g.order == 1) {
    interior = ni;
    edge = ne;
    return;
  }
  xcd_order(ni, interior);
  xcd_order(ne, edge);
}

// Arguments of stage `stage` (0..3) of an RK4 step of size dt; the tile
// pointer is set per group by the caller. `swap`: the mixed form's buffer
// roles of the step (y_n in the accumulator buffer on odd steps).
StageArgs stage_args(const sbp_system_s* s, double dt, int stage, int swap) {
  static const double A_next[3] = {0.5, 0.5, 1.0};                        // Rk4::A nonzero entries
  static const double Bw[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};  // Rk4::B
  static const double C[3] = {0.5, 0.5, 1.0};                             // Rk4::C
  double gamma;
  gamma_get(gamma);
  StageArgs a{};
  a.blocks = s->d_blocks;
  a.tiles = s->d_tiles;
  a.time = s->d_time;
  a.error = s->d_err;
  a.dt = dt;
  a.gamma = gamma;
  a.rk_b = Bw[stage];
  a.rk_a = stage < 3 ? A_next[stage] : 0.0;
  a.time_offset = stage == 0 ? 0.0 : dt * C[stage - 1];
  a.mixed = s->precision == SBP_MIXED;
  a.swap = a.mixed ? swap : 0;
  a.mixref = s->d_mixref;
  return a;
}

// Mixed form: bring y_n back into its own buffer after an odd number of
// steps, so that every host-side access finds the state in BUF_YN.
int normalize(sbp_system_s* s) {
  if (!s->parity) return SBP_OK;
  for (auto& h : s->blocks)
    SBP_TRY_HIP(launch_copy_planes(h.d_buf[BUF_ACC], h.ps, h.ld, h.d_buf[BUF_YN], h.ps, h.ld, 4, h.ny, h.nx, s->ctx->stream));
  s->parity = 0;
  return SBP_OK;
}

// One stage of one group (Eval edges get their boundary data at the stage
// time from a small kernel first). Schedule 0, split: the edge tiles, which
// evaluate the SATs of their block-edge points themselves, on the context's
// second stream, the interior tiles on `st`, concurrently; `st` continues
// once both are done (also under graph capture: a fork and a join).
// Schedule 1, merged: one dispatch of the edge instantiation over all tiles,
// the edge tiles first (it takes the interior path on interior tiles), on
// `st`. Schedule 2: the edge-SAT kernel evaluates the SATs of every block
// edge of the group, then the merged dispatch reads them.
hipError_t launch_group(const sbp_system_s* s, const Group& g, int shape, int sched, int mode, const StageArgs& a,
                        hipStream_t st) {
  const sbp_context_s* c = s->ctx;
  hipError_t e;
  // mixed form: the tiles' reference face terms, once per step (from y_n,
  // ahead of the first stage) or per right-hand-side launch
  if (a.mixed && (mode == MODE_S0 || mode == MODE_RHS)) {
    StageArgs ar = a;
    ar.tiles = s->d_tiles + g.etile_offset;
    if ((e = launch_mix_ref(ar, shape, mode, g.entiles + g.ntiles, st)) != hipSuccess) return e;
  }
  if (sched != 0) {
    if ((e = launch_eval_edges(a, s->d_evsegs + g.eval_off, g.neval, g.eval_maxlen, st)) != hipSuccess) return e;
    StageArgs am = a;
    if (sched == 2) {
      if ((e = launch_edge_sat(a, s->d_evsegs + g.sat_off, g.nsat, g.sat_maxlen, mode, st)) != hipSuccess) return e;
      am.sat_buf = 1;
    }
    am.tiles = s->d_tiles + g.etile_offset;
    return launch_stage(g.op_eta, g.op_xi, g.upwind, true, shape, mode, am, g.entiles + g.ntiles, st);
  }
  if ((e = hipEventRecord(c->ev_fork, st)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(c->aux, c->ev_fork, 0)) != hipSuccess) return e;
  hipError_t ee = launch_eval_edges(a, s->d_evsegs + g.eval_off, g.neval, g.eval_maxlen, c->aux);
  StageArgs ae = a;
  ae.tiles = s->d_tiles + g.etile_offset;
  if (ee == hipSuccess) ee = launch_stage(g.op_eta, g.op_xi, g.upwind, true, shape, mode, ae, g.entiles, c->aux);
  StageArgs ai = a;
  ai.tiles = s->d_tiles + g.tile_offset;
  // one of the two dispatches carries the time update of StageArgs::time_w
  if (g.entiles > 0)
    ai.time_w = nullptr;
  else
    ae.time_w = nullptr;
  const hipError_t ei = ee == hipSuccess ? launch_stage(g.op_eta, g.op_xi, g.upwind, false, shape, mode, ai, g.ntiles, st) : ee;
  // join even after an error, so that a capture in progress stays balanced
  if ((e = hipEventRecord(c->ev_join, c->aux)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(st, c->ev_join, 0)) != hipSuccess) return e;
  return ei;
}

// Autotuner timing of one group at one shape and schedule: a warm-up
// stage, then three rounds of the stage kernels that leave the state alone,
// between two events: modes S0, S1, S2 read y_n and write only stage inputs
// and increments; the mixed form's S3 too (it writes y_{n+1} into the other
// FP64 buffer, y_n's until the host swaps the roles), so its heaviest stage
// is timed with the others. (The FP64 S3 writes y_{n+1} over y_n.)
hipError_t time_group(const sbp_system_s* s, const Group& g, int shape, int sched, hipEvent_t e0, hipEvent_t e1,
                      float& ms) {
  hipStream_t st = s->ctx->stream;
  const StageArgs a = stage_args(s, 0.0, 0, 0);
  const int last = s->precision == SBP_MIXED ? MODE_S3 : MODE_S2;
  hipError_t e = launch_group(s, g, shape, sched, MODE_S0, a, st);
  if (e != hipSuccess) return e;
  if ((e = hipEventRecord(e0, st)) != hipSuccess) return e;
  for (int r = 0; r < 3; ++r)
    for (int m = MODE_S0; m <= last; ++m)
      if ((e = launch_group(s, g, shape, sched, m, a, st)) != hipSuccess) return e;
  if ((e = hipEventRecord(e1, st)) != hipSuccess) return e;
  if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
  return hipEventElapsedTime(&ms, e0, e1);
}

int rebuild(sbp_system_s* s) {
  if (!s->dirty) return SBP_OK;
  SBP_TRY(normalize(s));
  hipStream_t st = s->ctx->stream;
  // groups by operator pair and RHS kind
  s->groups.clear();
  for (int i = 0; i < (int)s->blocks.size(); ++i) {
    int e;
    bool up = block_upwind(s, s->blocks[i], &e);
    if (e) return e;
    bool found = false;
    for (auto& g : s->groups)
      if (g.op_eta == s->blocks[i].op_eta && g.op_xi == s->blocks[i].op_xi && g.upwind == up) {
        g.blocks.push_back(i);
        found = true;
      }
    if (!found) s->groups.push_back({s->blocks[i].op_eta, s->blocks[i].op_xi, up, {i}});
  }
  // tile shapes: forced, cached, or autotuned after the descriptors exist
  std::vector<TileDev> tiles;
  std::vector<BlockDev> devs(s->blocks.size());
  for (auto& g : s->groups) {
    int64_t pts = 0;
    for (int bi : g.blocks) pts += (int64_t)s->blocks[bi].nx * s->blocks[bi].ny;
    auto key = std::make_tuple(g.op_eta, g.op_xi, tune_kind(s, g), pts, (int)g.blocks.size());
    // cached choices: shape | schedule << 8, for groups tuned without a forced shape or schedule
    auto it = s->ctx->tuned.find(key);
    const bool cached =
        it != s->ctx->tuned.end() && s->forced_shape < 0 && s->forced_schedule < 0 && s->forced_order < 0;
    g.shape = s->forced_shape >= 0 ? s->forced_shape : cached ? (it->second & 0xff) : -1;
    g.sched = s->forced_schedule >= 0 ? s->forced_schedule : cached ? ((it->second >> 8) & 0xf) : -1;
    g.order = s->forced_order >= 0 ? s->forced_order : cached ? (it->second >> 12) : 0;
  }
  bool need_tune = false, need_sched = false;
  std::vector<bool> tune_cache(s->groups.size());
  for (size_t gi = 0; gi < s->groups.size(); ++gi) {
    const Group& g = s->groups[gi];
    need_tune = need_tune || g.shape < 0;
    need_sched = need_sched || g.sched < 0;
    tune_cache[gi] = g.shape < 0 && g.sched < 0 && s->forced_order < 0;
  }
  auto upload = [&](void) -> int {
    tiles.clear();
    for (auto& g : s->groups) {
      const int sh = g.shape < 0 ? 0 : g.shape;
      std::vector<TileDev> ti, te;
      build_tiles(s, g, sh, ti, te);
      g.etile_offset = (int)tiles.size();
      g.entiles = (int)te.size();
      tiles.insert(tiles.end(), te.begin(), te.end());
      g.tile_offset = (int)tiles.size();
      g.ntiles = (int)ti.size();
      tiles.insert(tiles.end(), ti.begin(), ti.end());
      for (int bi : g.blocks) devs[bi] = make_dev(s, bi, sh);
    }
    // per-tile slots (advance_adaptive's error sums): blocks in order
    s->total_tiles = 0;
    for (auto& d : devs) {
      d.tile_off = s->total_tiles;
      s->total_tiles += (int64_t)d.ntx * d.nty;
    }
    if (s->precision == SBP_MIXED && s->mixref_cap < s->total_tiles) {
      if (s->d_mixref) hipFree(s->d_mixref);
      s->d_mixref = nullptr;
      s->mixref_cap = 0;
      SBP_TRY_HIP(hipMalloc(&s->d_mixref, sizeof(double) * 16 * s->total_tiles));
      s->mixref_cap = s->total_tiles;
    }
    if (s->d_blocks) hipFree(s->d_blocks);
    if (s->d_tiles) hipFree(s->d_tiles);
    if (s->d_pieces) hipFree(s->d_pieces);
    if (s->d_edata) hipFree(s->d_edata);
    if (s->d_evsegs) hipFree(s->d_evsegs);
    if (s->d_progs) hipFree(s->d_progs);
    s->d_blocks = nullptr;
    s->d_tiles = nullptr;
    s->d_pieces = nullptr;
    s->d_edata = nullptr;
    s->d_evsegs = nullptr;
    s->d_progs = nullptr;
    // every block edge of each group in group order (SAT buffers of 4 x edge
    // length doubles), then the Eval edges of each group with their data
    // buffers
    std::vector<SatSeg> segs;
    std::vector<size_t> soff;
    size_t stotal = 0;
    for (auto& g : s->groups) {
      g.sat_off = (int)segs.size();
      g.sat_maxlen = 0;
      for (int bi : g.blocks)
        for (int d = 0; d < 4; ++d) {
          const int len = (d == SBP_NORTH || d == SBP_SOUTH) ? s->blocks[bi].nx : s->blocks[bi].ny;
          segs.push_back({(int32_t)bi, d});
          soff.push_back(stotal);
          stotal += 4 * (size_t)len;
          g.sat_maxlen = std::max(g.sat_maxlen, len);
        }
      g.nsat = (int)segs.size() - g.sat_off;
    }
    for (auto& g : s->groups) {
      g.eval_off = (int)segs.size();
      g.eval_maxlen = 0;
      for (int bi : g.blocks)
        for (int d = 0; d < 4; ++d)
          if (s->blocks[bi].bc_kind[d] == SBP_BC_EVAL) {
            const int len = (d == SBP_NORTH || d == SBP_SOUTH) ? s->blocks[bi].nx : s->blocks[bi].ny;
            segs.push_back({(int32_t)bi, d});
            soff.push_back(stotal);
            stotal += 4 * (size_t)len;
            g.eval_maxlen = std::max(g.eval_maxlen, len);
          }
      g.neval = (int)segs.size() - g.eval_off;
    }
    SBP_TRY_HIP(hipMalloc(&s->d_edata, sizeof(double) * std::max<size_t>(stotal, 1)));
    SBP_TRY_HIP(hipMalloc(&s->d_evsegs, sizeof(SatSeg) * std::max<size_t>(segs.size(), 1)));
    SBP_TRY_HIP(hipMemsetAsync(s->d_edata, 0, sizeof(double) * std::max<size_t>(stotal, 1), st));
    if (!segs.empty())
      SBP_TRY_HIP(hipMemcpyAsync(s->d_evsegs, segs.data(), sizeof(SatSeg) * segs.size(), hipMemcpyHostToDevice, st));
    for (auto& g : s->groups) {
      for (int k = g.sat_off; k < g.sat_off + g.nsat; ++k) devs[segs[k].block].sat[segs[k].dir] = s->d_edata + soff[k];
      for (int k = g.eval_off; k < g.eval_off + g.neval; ++k)
        devs[segs[k].block].edata[segs[k].dir] = s->d_edata + soff[k];
    }
    // Eval edges: their programs, 16-byte aligned, in one device array
    std::vector<unsigned char> progs;
    std::vector<std::pair<size_t, size_t>> pref;  // (block*4 + dir, offset)
    for (size_t bi = 0; bi < s->blocks.size(); ++bi)
      for (int d = 0; d < 4; ++d)
        if (s->blocks[bi].bc_kind[d] == SBP_BC_EVAL) {
          const auto& blob = *s->blocks[bi].bc_prog[d];
          pref.push_back({4 * bi + d, progs.size()});
          progs.insert(progs.end(), blob.begin(), blob.end());
          progs.resize((progs.size() + 15) / 16 * 16, 0);
        }
    if (!progs.empty()) {
      SBP_TRY_HIP(hipMalloc(&s->d_progs, progs.size()));
      SBP_TRY_HIP(hipMemcpyAsync(s->d_progs, progs.data(), progs.size(), hipMemcpyHostToDevice, st));
      for (auto& pr : pref)
        devs[pr.first / 4].eprog[pr.first % 4] = reinterpret_cast<const expr::Program*>(s->d_progs + pr.second);
    }
    // MultiGrid piece lists: one device array, each edge pointing into it
    std::vector<int> pieces;
    std::vector<size_t> poff(4 * s->blocks.size(), 0);
    for (size_t bi = 0; bi < s->blocks.size(); ++bi)
      for (int d = 0; d < 4; ++d) {
        poff[4 * bi + d] = pieces.size();
        pieces.insert(pieces.end(), s->blocks[bi].bc_pieces[d].begin(), s->blocks[bi].bc_pieces[d].end());
      }
    if (!pieces.empty()) {
      SBP_TRY_HIP(hipMalloc(&s->d_pieces, sizeof(int) * pieces.size()));
      SBP_TRY_HIP(hipMemcpyAsync(s->d_pieces, pieces.data(), sizeof(int) * pieces.size(), hipMemcpyHostToDevice, st));
      for (size_t bi = 0; bi < s->blocks.size(); ++bi)
        for (int d = 0; d < 4; ++d)
          if (!s->blocks[bi].bc_pieces[d].empty()) devs[bi].edge_pieces[d] = s->d_pieces + poff[4 * bi + d];
    }
    SBP_TRY_HIP(hipMalloc(&s->d_blocks, sizeof(BlockDev) * devs.size()));
    SBP_TRY_HIP(hipMalloc(&s->d_tiles, sizeof(TileDev) * std::max<size_t>(1, tiles.size())));
    SBP_TRY_HIP(hipMemcpyAsync(s->d_blocks, devs.data(), sizeof(BlockDev) * devs.size(), hipMemcpyHostToDevice, st));
    SBP_TRY_HIP(hipMemcpyAsync(s->d_tiles, tiles.data(), sizeof(TileDev) * tiles.size(), hipMemcpyHostToDevice, st));
    SBP_TRY_HIP(hipStreamSynchronize(st));
    return SBP_OK;
  };
  if (need_tune) {
    // Autotuner: time the stage-1..3 kernels (modes S0, S1, S2 read y_n and
    // write only the accumulator and the stage inputs, so the state is
    // untouched) for every configuration and keep the fastest per group.
    // two rounds over all configurations, each configuration's time the
    // smaller of its two (clock and cache state drift between timings)
    std::vector<int> best(s->groups.size(), 0);
    std::vector<float> best_ms(s->groups.size(), 1e30f);
    std::vector<std::vector<float>> cfg_ms(s->groups.size(), std::vector<float>(kNumTileShapes, 1e30f));
    hipEvent_t e0, e1;
    SBP_TRY_HIP(hipEventCreate(&e0));
    SBP_TRY_HIP(hipEventCreate(&e1));
    for (int round = 0; round < 2; ++round)
      for (int sh = 0; sh < kNumTileShapes; ++sh) {
        // temporarily give every untuned group this shape
        std::vector<int> saved;
        for (auto& g : s->groups) {
          saved.push_back(g.shape);
          if (g.shape < 0) g.shape = sh;
        }
        SBP_TRY(upload());
        for (size_t gi = 0; gi < s->groups.size(); ++gi) {
          Group& g = s->groups[gi];
          if (saved[gi] >= 0) continue;
          float ms = 0.f;
          const hipError_t w = time_group(s, g, sh, g.sched < 0 ? 0 : g.sched, e0, e1, ms);
          if (w == hipErrorNotSupported) continue;  // configuration not built for this operator pair
          SBP_TRY_HIP(w);
          cfg_ms[gi][sh] = std::min(cfg_ms[gi][sh], ms);
        }
        for (size_t gi = 0; gi < s->groups.size(); ++gi) s->groups[gi].shape = saved[gi];
      }
    // the three fastest configurations of each group again, three more
    // rounds each: shapes within a few percent of each other trade places
    // from one timing to the next, and the runs that follow last much longer
    constexpr int kFinalists = 3;
    std::vector<std::vector<int>> fin(s->groups.size());
    for (size_t gi = 0; gi < s->groups.size(); ++gi) {
      std::vector<int> order(kNumTileShapes);
      for (int sh = 0; sh < kNumTileShapes; ++sh) order[sh] = sh;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cfg_ms[gi][a] < cfg_ms[gi][b]; });
      for (int k = 0; k < kFinalists && k < kNumTileShapes; ++k)
        if (cfg_ms[gi][order[k]] < 1e30f) fin[gi].push_back(order[k]);
    }
    std::vector<std::vector<float>> fin_ms(s->groups.size(), std::vector<float>(kFinalists, 1e30f));
    for (int round = 0; round < 3; ++round)
      for (int k = 0; k < kFinalists; ++k) {
        std::vector<int> saved;
        bool any = false;
        for (size_t gi = 0; gi < s->groups.size(); ++gi) {
          Group& g = s->groups[gi];
          saved.push_back(g.shape);
          if (g.shape < 0 && k < (int)fin[gi].size()) {
            g.shape = fin[gi][k];
            any = true;
          }
        }
        if (any) {
          SBP_TRY(upload());
          for (size_t gi = 0; gi < s->groups.size(); ++gi) {
            Group& g = s->groups[gi];
            if (saved[gi] >= 0 || k >= (int)fin[gi].size()) continue;
            float ms = 0.f;
            SBP_TRY_HIP(time_group(s, g, g.shape, g.sched < 0 ? 0 : g.sched, e0, e1, ms));
            fin_ms[gi][k] = std::min(fin_ms[gi][k], ms);
          }
        }
        for (size_t gi = 0; gi < s->groups.size(); ++gi) s->groups[gi].shape = saved[gi];
      }
    for (size_t gi = 0; gi < s->groups.size(); ++gi)
      for (int k = 0; k < (int)fin[gi].size(); ++k)
        if (fin_ms[gi][k] < best_ms[gi]) {
          best_ms[gi] = fin_ms[gi][k];
          best[gi] = fin[gi][k];
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // the timing launches may have raised error bits on garbage accumulators: clear
    SBP_TRY_HIP(hipMemsetAsync(s->d_err, 0, sizeof(unsigned), st));
    for (size_t gi = 0; gi < s->groups.size(); ++gi)
      if (s->groups[gi].shape < 0) s->groups[gi].shape = best[gi];
  }
  if (need_sched) {
    // schedule and tile order at the chosen shape: each of the three
    // schedules in each of the two orders, four rounds each, the smaller time
    // of each (they differ by a few percent at most)
    hipEvent_t e0, e1;
    SBP_TRY_HIP(hipEventCreate(&e0));
    SBP_TRY_HIP(hipEventCreate(&e1));
    std::vector<std::array<std::array<float, 3>, 2>> t(s->groups.size());
    for (auto& tg : t)
      for (auto& to : tg) to.fill(1e30f);
    for (int o = 0; o < 2; ++o) {
      if (s->forced_order >= 0 && o != s->forced_order) continue;
      for (auto& g : s->groups)
        if (g.sched < 0) g.order = o;
      SBP_TRY(upload());
      for (size_t gi = 0; gi < s->groups.size(); ++gi) {
        Group& g = s->groups[gi];
        if (g.sched >= 0) continue;
        for (int round = 0; round < 4; ++round)
          for (int m = 0; m < 3; ++m) {
            float ms = 0.f;
            SBP_TRY_HIP(time_group(s, g, g.shape, m, e0, e1, ms));
            t[gi][o][m] = std::min(t[gi][o][m], ms);
          }
      }
    }
    for (size_t gi = 0; gi < s->groups.size(); ++gi) {
      Group& g = s->groups[gi];
      if (g.sched >= 0) continue;
      int bo = 0, bm = 0;
      for (int o = 0; o < 2; ++o)
        for (int m = 0; m < 3; ++m)
          if (t[gi][o][m] < t[gi][bo][bm]) {
            bo = o;
            bm = m;
          }
      g.order = bo;
      g.sched = bm;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    SBP_TRY_HIP(hipMemsetAsync(s->d_err, 0, sizeof(unsigned), st));
  }
  for (size_t gi = 0; gi < s->groups.size(); ++gi) {
    const Group& g = s->groups[gi];
    if (!tune_cache[gi]) continue;
    int64_t pts = 0;
    for (int bi : g.blocks) pts += (int64_t)s->blocks[bi].nx * s->blocks[bi].ny;
    s->ctx->tuned[std::make_tuple(g.op_eta, g.op_xi, tune_kind(s, g), pts, (int)g.blocks.size())] =
        g.shape | (g.sched << 8) | (g.order << 12);
  }
  SBP_TRY(upload());
  destroy_graph(s);
  s->dirty = false;
  return SBP_OK;
}

int check_err(sbp_system_s* s) {
  unsigned w = 0;
  SBP_TRY_HIP(hipMemcpyAsync(&w, s->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  if (w) SBP_TRY_HIP(hipMemsetAsync(s->d_err, 0, sizeof(unsigned), s->ctx->stream));
  return status_from_errword(w);
}

// One fused RK4 step: four stage launches per group, the time update
// (integrate::<Rk4>, utils/integrate/src/lib.rs:176-222) carried by the first
// two: stage S0 copies the step's start time d_time[0] into d_time[2], which
// stages S1-S3 read, and S1 writes d_time[2] + dt into d_time[0] (no
// dispatch of a stage reads the slot that stage writes).
int enqueue_step(sbp_system_s* s, double dt, int swap, hipStream_t st) {
  double* const t_step = s->d_time;
  double* const t_stage = s->d_time + 2;
  for (int stage = 0; stage < 4; ++stage) {
    StageArgs a = stage_args(s, dt, stage, swap);
    if (stage > 0) a.time = t_stage;
    for (size_t gi = 0; gi < s->groups.size(); ++gi) {
      StageArgs ag = a;
      if (gi == 0 && stage < 2) ag.time_w = stage == 0 ? t_stage : t_step;
      SBP_TRY_HIP(launch_group(s, s->groups[gi], s->groups[gi].shape, s->groups[gi].sched, stage, ag, st));
    }
  }
  return SBP_OK;
}

int ensure_graph(sbp_system_s* s, double dt, int steps) {
  if (s->graph && s->graph_dt == dt && s->graph_steps == steps) return SBP_OK;
  destroy_graph(s);
  hipStream_t st = s->ctx->stream;
  SBP_TRY_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = SBP_OK;
  // the mixed form alternates buffer roles; its graphs hold an even number of steps
  for (int i = 0; i < steps && rc == SBP_OK; ++i) rc = enqueue_step(s, dt, i & 1, st);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != SBP_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  SBP_TRY_HIP(e);
  s->graph_def = g;
  SBP_TRY_HIP(hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0));
  s->graph_dt = dt;
  s->graph_steps = steps;
  return SBP_OK;
}

// Steps per captured graph: small systems are launch-bound, so they batch
// several steps into one graph.
int steps_per_graph(const sbp_system_s* s) {
  int64_t pts = 0;
  for (auto& b : s->blocks) pts += (int64_t)b.nx * b.ny;
  if (pts >= (int64_t)1 << 22) return s->precision == SBP_MIXED ? 2 : 1;
  if (pts >= (int64_t)1 << 18) return 4;
  return 16;
}

int advance_impl(sbp_system_s* s, double dt, int64_t nsteps) {
  SBP_TRY(rebuild(s));
  const bool mixed = s->precision == SBP_MIXED;
  const int G = steps_per_graph(s);
  int64_t done = 0;
  // graphs start with y_n in its own buffer
  if (s->parity && nsteps > 0) {
    SBP_TRY(enqueue_step(s, dt, s->parity, s->ctx->stream));
    s->parity ^= 1;
    ++done;
  }
  if (nsteps - done >= G) {
    SBP_TRY(ensure_graph(s, dt, G));
    for (; done + G <= nsteps; done += G) SBP_TRY_HIP(hipGraphLaunch(s->graph, s->ctx->stream));
  }
  for (; done < nsteps; ++done) {
    SBP_TRY(enqueue_step(s, dt, s->parity, s->ctx->stream));
    if (mixed) s->parity ^= 1;
  }
  return SBP_OK;
}

int ensure_partials(sbp_system_s* s) {
  if (s->d_partials) return SBP_OK;
  SBP_TRY_HIP(hipMalloc(&s->d_partials, sizeof(double) * kReduceScratch));
  return SBP_OK;
}

// At least `sets` Runge-Kutta work sets per block (zeroed once); descriptors
// are rebuilt when a block's sets grow.
int ensure_work_sets(sbp_system_s* s, int sets) {
  hipStream_t st = s->ctx->stream;
  for (auto& h : s->blocks) {
    if (h.rk_sets >= sets) continue;
    SBP_TRY_HIP(hipStreamSynchronize(st));
    if (h.d_rk) hipFree(h.d_rk);
    h.d_rk = nullptr;
    h.rk_sets = 0;
    SBP_TRY_HIP(hipMalloc(&h.d_rk, sizeof(double) * sets * 4 * h.ps));
    SBP_TRY_HIP(hipMemsetAsync(h.d_rk, 0, sizeof(double) * sets * 4 * h.ps, st));
    h.rk_sets = sets;
    s->dirty = true;
  }
  return SBP_OK;
}

// Reads the result pair of a reduction in `scratch` (waits for the stream).
int reduce_result(const double* scratch, hipStream_t st, double out[2]) {
  SBP_TRY_HIP(hipMemcpyAsync(out, scratch + 2 * kReduceMaxWG, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  SBP_TRY_HIP(hipStreamSynchronize(st));
  return SBP_OK;
}

int h2_sum(sbp_system_s* s, int bi, const double* b, const VortexDev* vp_host, double t, long double& sum) {
  SBP_TRY(normalize(s));
  HostBlock& h = s->blocks[bi];
  SBP_TRY(ensure_partials(s));
  hipStream_t st = s->ctx->stream;
  double gamma;
  gamma_get(gamma);
  H2Args R{};
  R.a = h.d_buf[BUF_YN];
  R.psa = h.ps;
  R.lda = h.ld;
  R.b = b;
  R.psb = h.ps;
  R.ldb = h.ld;
  if (vp_host) R.vp = *vp_host;
  R.use_vortex = vp_host != nullptr;
  R.gamma = gamma;
  R.t = t;
  R.xc = h.d_x;
  R.yc = h.d_y;
  R.ldc = h.ld;
  R.ny = h.ny;
  R.nx = h.nx;
  SBP_TRY_HIP(launch_h2_sum(h.op_eta, h.op_xi, R, s->d_partials, st));
  double r[2];
  SBP_TRY(reduce_result(s->d_partials, st, r));
  sum = (long double)r[0] + (long double)r[1];
  return SBP_OK;
}

}  // namespace

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

const char* sbp_status_string(int status) {
  switch (status) {
    case SBP_OK: return "ok";
    case SBP_ERR_INVALID: return "invalid argument";
    case SBP_ERR_SHAPE: return "operator needs n >= 2M and n >= N";
    case SBP_ERR_DETJ: return "grid has det J <= 0";
    case SBP_ERR_RHO: return "density <= 0";
    case SBP_ERR_PRESSURE: return "pressure <= 0";
    case SBP_ERR_GAMMA: return "GAMMA is not defined";
    case SBP_ERR_EDGE: return "neighbouring edges differ in length";
    case SBP_ERR_HIP: return "HIP runtime error";
    case SBP_ERR_NOT_UPWIND: return "This is not an upwind operator";
    case SBP_ERR_OOM: return "device out of memory";
    case SBP_ERR_GAMMA_SET: return "GAMMA already set";
    case SBP_ERR_SIZE: return "size beyond the entry point's maximum";
    case SBP_ERR_EXPR: return "expression does not compile or evaluate";
    default: return "unknown status";
  }
}

int sbp_set_gamma(double gamma) {
  std::lock_guard<std::mutex> lk(g_gamma_mu);
  if (g_gamma_set) return gamma == g_gamma ? SBP_OK : SBP_ERR_GAMMA_SET;
  if (!(gamma > 1.0)) return SBP_ERR_INVALID;
  g_gamma = gamma;
  g_gamma_set = true;
  return SBP_OK;
}

int sbp_get_gamma(double* gamma) {
  if (!gamma) return SBP_ERR_INVALID;
  double g;
  if (!gamma_get(g)) return SBP_ERR_GAMMA;
  *gamma = g;
  return SBP_OK;
}

int sbp_operator_info(int op, int* M, int* N, int* D, int* is_h2, int* is_upwind) {
  OpInfo o;
  if (!op_info(op, o)) return SBP_ERR_INVALID;
  if (M) *M = o.M;
  if (N) *N = o.N;
  if (D) *D = o.D;
  if (is_h2) *is_h2 = o.h2;
  if (is_upwind) *is_upwind = o.upwind;
  return SBP_OK;
}

int sbp_operator_h(int op, double* h_out) {
  OpInfo o;
  if (!op_info(op, o) || !h_out) return SBP_ERR_INVALID;
  for (int i = 0; i < o.M; ++i) h_out[i] = o.h[i];
  return SBP_OK;
}

int sbp_operator_coefficients(int op, int kind, double* start, double* diag, double* end) {
  if (!start || !diag || !end || (kind != SBP_DIFF && kind != SBP_DISS)) return SBP_ERR_INVALID;
  bool ok = false;
  switch (op) {
    case UPWIND4: coefs_of<ops::Upwind4>(kind, start, diag, end, ok); break;
    case UPWIND4H2: coefs_of<ops::Upwind4h2>(kind, start, diag, end, ok); break;
    case UPWIND9: coefs_of<ops::Upwind9>(kind, start, diag, end, ok); break;
    case UPWIND9H2: coefs_of<ops::Upwind9h2>(kind, start, diag, end, ok); break;
    case SBP4: coefs_of<ops::SBP4>(kind, start, diag, end, ok); break;
    case SBP8: coefs_of<ops::SBP8>(kind, start, diag, end, ok); break;
    default: return SBP_ERR_INVALID;
  }
  return ok ? SBP_OK : SBP_ERR_NOT_UPWIND;
}

int sbp_context_create(int device, sbp_context* out) {
  if (!out) return SBP_ERR_INVALID;
  int n = 0;
  SBP_TRY_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(device));
  auto* c = new sbp_context_s();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return from_hip(e);
  }
  c->own_stream = true;
  if ((e = hipMalloc(&c->d_err, sizeof(unsigned))) != hipSuccess || (e = hipMemset(c->d_err, 0, sizeof(unsigned))) != hipSuccess ||
      (e = hipMalloc(&c->d_tmpblock, sizeof(BlockDev))) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess ||
      (e = noalloc_scratch_of(device, &c->noalloc_scratch)) != hipSuccess) {
    if (c->aux) hipStreamDestroy(c->aux);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    hipStreamDestroy(c->stream);
    delete c;
    return from_hip(e);
  }
  *out = c;
  return SBP_OK;
}

int sbp_context_destroy(sbp_context ctx) {
  if (!ctx) return SBP_ERR_INVALID;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipStreamSynchronize(ctx->aux);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  hipStreamDestroy(ctx->aux);
  hipEventDestroy(ctx->ev_fork);
  hipEventDestroy(ctx->ev_join);
  hipFree(ctx->d_err);
  hipFree(ctx->d_tmpblock);
  if (ctx->d_tmptiles) hipFree(ctx->d_tmptiles);
  delete ctx;
  return SBP_OK;
}

int sbp_context_set_stream(sbp_context ctx, void* hip_stream) {
  if (!ctx) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  SBP_TRY_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  if (hip_stream) {
    ctx->stream = (hipStream_t)hip_stream;
    ctx->own_stream = false;
  } else {
    SBP_TRY_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  return SBP_OK;
}

int sbp_context_synchronize(sbp_context ctx) {
  if (!ctx) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return SBP_OK;
}

static int apply(sbp_context ctx, int op, int kind, int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny,
                 int nx) {
  if (!ctx || !prev || !fut || ny <= 0 || nx <= 0 || ldp < nx || ldf < nx) return SBP_ERR_INVALID;
  OpInfo o;
  if (!op_info(op, o)) return SBP_ERR_INVALID;
  if (kind == SBP_DISS && !o.upwind) return SBP_ERR_NOT_UPWIND;
  const int n = axis == 0 ? nx : ny;
  if (n < 2 * o.M || n < o.N) return SBP_ERR_SHAPE;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  SBP_TRY_HIP(launch_apply_op(op, kind, axis, prev, ldp, fut, ldf, ny, nx, ctx->stream));
  return SBP_OK;
}

int sbp_diffxi(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx) {
  return apply(ctx, op, SBP_DIFF, 0, prev, ld_prev, fut, ld_fut, ny, nx);
}
int sbp_diffeta(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx) {
  return apply(ctx, op, SBP_DIFF, 1, prev, ld_prev, fut, ld_fut, ny, nx);
}
int sbp_dissxi(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx) {
  return apply(ctx, op, SBP_DISS, 0, prev, ld_prev, fut, ld_fut, ny, nx);
}
int sbp_disseta(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx) {
  return apply(ctx, op, SBP_DISS, 1, prev, ld_prev, fut, ld_fut, ny, nx);
}

int sbp_interpolate(sbp_context ctx, int interp_op, int fine2coarse, const double* in, int nin, double* out, int nout) {
  if (interp_op < 0 || interp_op >= INTERP_COUNT) return SBP_ERR_INVALID;
  // lengths first (a host-only check, also without a context)
  const bool f2c = fine2coarse != 0;
  if (!interp::coupling_valid(interp_op, f2c, f2c ? nout : nin, f2c ? nin : nout)) return SBP_ERR_SHAPE;
  if (!ctx || !in || !out) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  SBP_TRY_HIP(launch_interpolate(interp_op, f2c, in, nin, out, nout, ctx->stream));
  return SBP_OK;
}

int sbp_interp_info(int interp_op, int fine2coarse, int* shape6) {
  if (!shape6 || interp_op < 0 || interp_op >= INTERP_COUNT) return SBP_ERR_INVALID;
  const interp::Table T = interp::table(interp_op, fine2coarse != 0);
  const int v[6] = {T.brows, T.bcols, T.drows, T.dcols, T.jump0, T.jump1};
  for (int i = 0; i < 6; ++i) shape6[i] = v[i];
  return SBP_OK;
}

int sbp_interp_coefficients(int interp_op, int fine2coarse, double* block, double* diag) {
  if (!block || !diag || interp_op < 0 || interp_op >= INTERP_COUNT) return SBP_ERR_INVALID;
  const interp::Table T = interp::table(interp_op, fine2coarse != 0);
  for (int i = 0; i < T.brows * T.bcols; ++i) block[i] = T.block[i];
  for (int i = 0; i < T.drows * T.dcols; ++i) diag[i] = T.diag[i];
  return SBP_OK;
}

static int check_shape(int op_eta, int op_xi, int ny, int nx) {
  OpInfo oe, ox;
  if (!op_info(op_eta, oe) || !op_info(op_xi, ox)) return SBP_ERR_INVALID;
  if (nx < 2 * ox.M || nx < ox.N || ny < 2 * oe.M || ny < oe.N) return SBP_ERR_SHAPE;
  return SBP_OK;
}

int sbp_metrics(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* x, const double* y, double* metrics) {
  if (!ctx || !x || !y || !metrics) return SBP_ERR_INVALID;
  SBP_TRY(check_shape(op_eta, op_xi, ny, nx));
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  const int64_t ps = (int64_t)nx * ny;
  double* tmp = nullptr;
  SBP_TRY_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * 4 * ps, ctx->stream));
  hipStream_t st = ctx->stream;
  int rc = SBP_OK;
  if (rc == SBP_OK) rc = from_hip(launch_apply_op(op_xi, SBP_DIFF, 0, x, nx, tmp, nx, ny, nx, st));
  if (rc == SBP_OK) rc = from_hip(launch_apply_op(op_eta, SBP_DIFF, 1, x, nx, tmp + ps, nx, ny, nx, st));
  if (rc == SBP_OK) rc = from_hip(launch_apply_op(op_xi, SBP_DIFF, 0, y, nx, tmp + 2 * ps, nx, ny, nx, st));
  if (rc == SBP_OK) rc = from_hip(launch_apply_op(op_eta, SBP_DIFF, 1, y, nx, tmp + 3 * ps, nx, ny, nx, st));
  if (rc == SBP_OK)
    rc = from_hip(launch_metrics_finish(tmp, tmp + ps, tmp + 2 * ps, tmp + 3 * ps, metrics, ps, nx, ny, nx, ctx->d_err, st));
  hipFreeAsync(tmp, st);
  if (rc != SBP_OK) return rc;
  unsigned w = 0;
  SBP_TRY_HIP(hipMemcpyAsync(&w, ctx->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  SBP_TRY_HIP(hipStreamSynchronize(st));
  if (w) SBP_TRY_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(unsigned), st));
  return status_from_errword(w);
}

// The context's sticky error word (device errors of operator-level calls),
// read and cleared.
static int ctx_check_err(sbp_context ctx) {
  unsigned w = 0;
  SBP_TRY_HIP(hipMemcpyAsync(&w, ctx->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(ctx->stream));
  if (w) SBP_TRY_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(unsigned), ctx->stream));
  return status_from_errword(w);
}

// One block's right-hand side on caller buffers (contiguous (4, ny, nx)
// fields, (5, ny, nx) metrics) with the upwind terms when `up`, the SATs of
// the caller's boundary data when `boundaries` is given and none otherwise
// (RHS_no_SAT).
static int rhs_block(sbp_context ctx, bool up, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
                     const double* metrics, const double* const* boundaries) {
  SBP_TRY(check_shape(op_eta, op_xi, ny, nx));
  // contiguous planes (pitch nx) with the stage kernels' 32-bit offsets
  if (5 * (int64_t)nx * ny * 8 >= ((int64_t)1 << 32)) return SBP_ERR_SIZE;
  OpInfo oe, ox;
  op_info(op_eta, oe);
  op_info(op_xi, ox);
  if (up && !(oe.upwind && ox.upwind)) return SBP_ERR_NOT_UPWIND;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  if (boundaries)
    for (int d = 0; d < 4; ++d)
      if (!boundaries[d]) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  const int shape = 0;
  const TileShape& ts = kTileShapes[shape];
  BlockDev b{};
  b.nx = nx;
  b.ny = ny;
  b.ld = nx;
  b.ps = (int64_t)nx * ny;
  b.op_xi = op_xi;
  b.op_eta = op_eta;
  b.upwind = up;
  b.dt_scale = 1.0;
  b.ntx = ntiles_along(nx, ts.tx);
  b.nty = ntiles_along(ny, ts.ty);
  b.metrics = metrics;
  b.rhs_in = const_cast<double*>(y);
  b.rhs_out = k;
  b.idx_xi = 1.0 / (ox.h2 ? 1.0 / (double)(nx - 2) : 1.0 / (double)(nx - 1));
  b.idx_eta = 1.0 / (oe.h2 ? 1.0 / (double)(ny - 2) : 1.0 / (double)(ny - 1));
  b.hi_xi = (ox.h2 ? (double)(nx - 2) : (double)(nx - 1)) / ox.h[0];
  b.hi_eta = (oe.h2 ? (double)(ny - 2) : (double)(ny - 1)) / oe.h[0];
  for (int d = 0; d < 4; ++d) {
    b.edge_src[d] = boundaries ? EDGE_EXPLICIT : EDGE_NONE;
    b.edge_ptr[d] = boundaries ? boundaries[d] : nullptr;
    b.edge_cs[d] = (d == SBP_NORTH || d == SBP_SOUTH) ? nx : ny;
  }
  std::vector<TileDev> tiles;
  for (int ty = 0; ty < b.nty; ++ty)
    for (int tx = 0; tx < b.ntx; ++tx) tiles.push_back({0, (int16_t)tx, (int16_t)ty});
  hipStream_t st = ctx->stream;
  SBP_TRY_HIP(hipStreamSynchronize(st));
  // the stage kernel reads the metric pairs; build them from the caller's planes
  double* pairs = nullptr;
  SBP_TRY_HIP(hipMalloc((void**)&pairs, sizeof(double) * 4 * b.ps));
  struct FreeOnExit {
    double* p;
    ~FreeOnExit() { hipFree(p); }
  } free_pairs{pairs};
  SBP_TRY_HIP(launch_metric_pairs(metrics, b.ps, b.ld, ny, nx, pairs, st));
  b.mpair = pairs;
  if (ctx->tmptiles_cap < (int)tiles.size()) {
    if (ctx->d_tmptiles) hipFree(ctx->d_tmptiles);
    ctx->d_tmptiles = nullptr;
    SBP_TRY_HIP(hipMalloc(&ctx->d_tmptiles, sizeof(TileDev) * tiles.size()));
    ctx->tmptiles_cap = (int)tiles.size();
  }
  SBP_TRY_HIP(hipMemcpy(ctx->d_tmpblock, &b, sizeof(BlockDev), hipMemcpyHostToDevice));
  SBP_TRY_HIP(hipMemcpy(ctx->d_tmptiles, tiles.data(), sizeof(TileDev) * tiles.size(), hipMemcpyHostToDevice));
  StageArgs a{};
  a.blocks = ctx->d_tmpblock;
  a.tiles = ctx->d_tmptiles;
  a.time = nullptr;
  a.error = ctx->d_err;
  a.gamma = gamma;
  SBP_TRY_HIP(launch_stage(op_eta, op_xi, up, true, shape, MODE_RHS, a, (int)tiles.size(), st));
  return ctx_check_err(ctx);
}

int sbp_rhs(sbp_context ctx, int kind, int op_eta, int op_xi, int ny, int nx, double* k, const double* y, const double* metrics,
            const double* const* boundaries) {
  if (!ctx || !k || !y || !metrics || !boundaries) return SBP_ERR_INVALID;
  if (kind != SBP_RHS_TRAD && kind != SBP_RHS_UPWIND) return SBP_ERR_INVALID;
  return rhs_block(ctx, kind == SBP_RHS_UPWIND, op_eta, op_xi, ny, nx, k, y, metrics, boundaries);
}

int sbp_rhs_no_sat(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, double* k, const double* y, const double* metrics) {
  if (!ctx || !k || !y || !metrics) return SBP_ERR_INVALID;
  OpInfo oe, ox;
  if (!op_info(op_eta, oe) || !op_info(op_xi, ox)) return SBP_ERR_INVALID;
  // op.upwind().is_some() for the operator pair (sbp/src/operators.rs:137-150)
  return rhs_block(ctx, oe.upwind && ox.upwind, op_eta, op_xi, ny, nx, k, y, metrics, nullptr);
}

// SAT of the edges in `dirs` (in that order) added to k on caller buffers.
static int sat_edges(sbp_context ctx, const int* dirs, int ndirs, int op_eta, int op_xi, int ny, int nx, double* k,
                     const double* y, const double* metrics, const double* const* g) {
  if (!ctx || !k || !y || !metrics || !g) return SBP_ERR_INVALID;
  SBP_TRY(check_shape(op_eta, op_xi, ny, nx));
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  OpInfo oe, ox;
  op_info(op_eta, oe);
  op_info(op_xi, ox);
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  for (int i = 0; i < ndirs; ++i) {
    if (!g[i]) return SBP_ERR_INVALID;
    SatEdgeArgs a{};
    a.dir = dirs[i];
    a.k = k;
    a.y = y;
    a.met = metrics;
    a.g = g[i];
    a.ps = (int64_t)nx * ny;
    a.ld = nx;
    a.ny = ny;
    a.nx = nx;
    const bool ns = a.dir == SBP_NORTH || a.dir == SBP_SOUTH;
    // hi = (n - 1)/H[0], (n - 2)/H[0] on h2 grids (lib.rs:1034-1038, :1095-1099)
    a.hi = ns ? (oe.h2 ? (double)(ny - 2) : (double)(ny - 1)) / oe.h[0]
              : (ox.h2 ? (double)(nx - 2) : (double)(nx - 1)) / ox.h[0];
    a.gamma = gamma;
    SBP_TRY_HIP(launch_sat_edge(a, ctx->stream));
  }
  SBP_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return SBP_OK;
}

int sbp_sat_characteristics(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
                            const double* metrics, const double* const* boundaries) {
  static const int order[4] = {SBP_NORTH, SBP_SOUTH, SBP_EAST, SBP_WEST};
  if (!boundaries) return SBP_ERR_INVALID;
  return sat_edges(ctx, order, 4, op_eta, op_xi, ny, nx, k, y, metrics, boundaries);
}

int sbp_sat_edge(sbp_context ctx, int dir, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
                 const double* metrics, const double* boundary) {
  if (dir < 0 || dir > 3) return SBP_ERR_INVALID;
  const double* g[1] = {boundary};
  return sat_edges(ctx, &dir, 1, op_eta, op_xi, ny, nx, k, y, metrics, g);
}

// Field::h2_err of two fields on caller buffers through `scratch`
// (kReduceScratch doubles of device memory).
static int h2_err_with(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b,
                       double* scratch, double* out) {
  H2Args R{};
  R.a = a;
  R.psa = (int64_t)nx * ny;
  R.lda = nx;
  R.b = b;
  R.psb = R.psa;
  R.ldb = nx;
  R.ny = ny;
  R.nx = nx;
  SBP_TRY_HIP(launch_h2_sum(op_eta, op_xi, R, scratch, ctx->stream));
  double r[2];
  SBP_TRY(reduce_result(scratch, ctx->stream, r));
  *out = std::sqrt(r[0] + r[1]);
  return SBP_OK;
}

// Sizes first (host-only checks, also without a context), then handles.
static int h2_err_args(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b,
                       double* out) {
  OpInfo oe, ox;
  if (!op_info(op_eta, oe) || !op_info(op_xi, ox)) return SBP_ERR_INVALID;
  if (nx <= 0 || ny <= 0 || (int64_t)nx * ny > (int64_t)INT32_MAX) return SBP_ERR_SIZE;
  // the weights [H, 1.., rev(H)] need n >= 2M (lib.rs:389-395)
  if (nx < 2 * ox.M || ny < 2 * oe.M) return SBP_ERR_SHAPE;
  if (!ctx || !a || !b || !out) return SBP_ERR_INVALID;
  return SBP_OK;
}

int sbp_h2_err(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b, double* out) {
  SBP_TRY(h2_err_args(ctx, op_eta, op_xi, ny, nx, a, b, out));
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  double* scratch = nullptr;
  SBP_TRY_HIP(hipMallocAsync((void**)&scratch, sizeof(double) * kReduceScratch, ctx->stream));
  const int rc = h2_err_with(ctx, op_eta, op_xi, ny, nx, a, b, scratch, out);
  hipFreeAsync(scratch, ctx->stream);
  return rc;
}

int sbp_h2_err_noalloc(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b,
                       double* out) {
  SBP_TRY(h2_err_args(ctx, op_eta, op_xi, ny, nx, a, b, out));
  if (!ctx->noalloc_scratch) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  // every partial the finish pass reads is written by this call's partial
  // pass first (stream order), so nothing in the scratch needs clearing
  return h2_err_with(ctx, op_eta, op_xi, ny, nx, a, b, ctx->noalloc_scratch, out);
}

int sbp_h2linspace(double start, double end, int n, double* out) {
  // the reference's arithmetic, two roundings per point (Rust never contracts)
#pragma clang fp contract(off)
  if (!out || n < 2) return SBP_ERR_INVALID;
  const double h = (end - start) / (double)(n - 2);
  for (int i = 0; i < n; ++i) out[i] = i == 0 ? start : i == n - 1 ? end : start + h * ((double)i - 0.5);
  return SBP_OK;
}

int sbp_system_create(sbp_context ctx, int nblocks, const int* ny, const int* nx, const int* op_eta, const int* op_xi,
                      const double* const* x, const double* const* y, sbp_system* out) {
  if (nblocks <= 0 || nblocks > 32767 || !ny || !nx || !op_eta || !op_xi) return SBP_ERR_INVALID;
  // sizes first (host-side checks, before the handles)
  for (int i = 0; i < nblocks; ++i) {
    SBP_TRY(check_shape(op_eta[i], op_xi[i], ny[i], nx[i]));
    if (nx[i] > 32767 * 16 || ny[i] > 32767 * 16 || !planes_fit(nx[i], ny[i])) return SBP_ERR_SIZE;
  }
  if (!ctx || !x || !y || !out) return SBP_ERR_INVALID;
  for (int i = 0; i < nblocks; ++i)
    if (!x[i] || !y[i]) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<sbp_system_s> s(new sbp_system_s());
  s->ctx = ctx;
  hipStream_t st = ctx->stream;
  auto fail = [&](int rc) {
    sbp_system_destroy(s.release());
    return rc;
  };
  s->blocks.resize(nblocks);
  for (int i = 0; i < nblocks; ++i) {
    HostBlock& h = s->blocks[i];
    h.nx = nx[i];
    h.ny = ny[i];
    h.op_eta = op_eta[i];
    h.op_xi = op_xi[i];
    h.ld = pitch_of(h.nx);
    h.ps = h.ld * h.ny;
    std::memset(h.bc_vortex, 0, sizeof(h.bc_vortex));
  }
  int rc = SBP_OK;
  for (int i = 0; i < nblocks && rc == SBP_OK; ++i) {
    HostBlock& h = s->blocks[i];
    const size_t plane = sizeof(double) * (size_t)h.ps;
    if ((rc = from_hip(hipMalloc((void**)&h.d_metrics, 5 * plane))) != SBP_OK) break;
    if ((rc = from_hip(hipMalloc((void**)&h.d_mpair, 4 * plane))) != SBP_OK) break;
    if ((rc = from_hip(hipMalloc((void**)&h.d_x, plane))) != SBP_OK) break;
    if ((rc = from_hip(hipMalloc((void**)&h.d_y, plane))) != SBP_OK) break;
    for (int r = 0; r < 4 && rc == SBP_OK; ++r) rc = from_hip(hipMalloc((void**)&h.d_buf[r], 4 * plane));
    if (rc != SBP_OK) break;
    // zero everything once so that pitch padding never holds garbage
    for (int r = 0; r < 4 && rc == SBP_OK; ++r) rc = from_hip(hipMemsetAsync(h.d_buf[r], 0, 4 * plane, st));
    if (rc == SBP_OK) rc = from_hip(hipMemsetAsync(h.d_metrics, 0, 5 * plane, st));
    if (rc == SBP_OK) rc = from_hip(hipMemsetAsync(h.d_mpair, 0, 4 * plane, st));
    if (rc == SBP_OK)
      rc = from_hip(hipMemcpy2DAsync(h.d_x, h.ld * 8, x[i], (size_t)h.nx * 8, (size_t)h.nx * 8, h.ny, hipMemcpyHostToDevice, st));
    if (rc == SBP_OK)
      rc = from_hip(hipMemcpy2DAsync(h.d_y, h.ld * 8, y[i], (size_t)h.nx * 8, (size_t)h.nx * 8, h.ny, hipMemcpyHostToDevice, st));
    // metrics (grid.rs:95-139); the accumulator buffer holds the four derivatives
    double* t = h.d_buf[BUF_ACC];
    if (rc == SBP_OK) rc = from_hip(launch_apply_op(h.op_xi, SBP_DIFF, 0, h.d_x, h.ld, t, h.ld, h.ny, h.nx, st));
    if (rc == SBP_OK) rc = from_hip(launch_apply_op(h.op_eta, SBP_DIFF, 1, h.d_x, h.ld, t + h.ps, h.ld, h.ny, h.nx, st));
    if (rc == SBP_OK) rc = from_hip(launch_apply_op(h.op_xi, SBP_DIFF, 0, h.d_y, h.ld, t + 2 * h.ps, h.ld, h.ny, h.nx, st));
    if (rc == SBP_OK) rc = from_hip(launch_apply_op(h.op_eta, SBP_DIFF, 1, h.d_y, h.ld, t + 3 * h.ps, h.ld, h.ny, h.nx, st));
    if (rc == SBP_OK)
      rc = from_hip(launch_metrics_finish(t, t + h.ps, t + 2 * h.ps, t + 3 * h.ps, h.d_metrics, h.ps, h.ld, h.ny, h.nx, ctx->d_err,
                                          st));
    if (rc == SBP_OK) rc = from_hip(launch_metric_pairs(h.d_metrics, h.ps, h.ld, h.ny, h.nx, h.d_mpair, st));
  }
  if (rc != SBP_OK) return fail(rc);
  // [0] the system time, [1] the explicit time of sbp_system_rhs, [2] the
  // stage time base inside an RK4 step (enqueue_step)
  if ((rc = from_hip(hipMalloc(&s->d_time, sizeof(double) * 3))) != SBP_OK) return fail(rc);
  s->d_time_rhs = s->d_time + 1;
  if ((rc = from_hip(hipMemsetAsync(s->d_time, 0, sizeof(double) * 3, st))) != SBP_OK) return fail(rc);
  if ((rc = from_hip(hipMalloc(&s->d_err, sizeof(unsigned)))) != SBP_OK) return fail(rc);
  if ((rc = from_hip(hipMemsetAsync(s->d_err, 0, sizeof(unsigned), st))) != SBP_OK) return fail(rc);
  unsigned w = 0;
  if ((rc = from_hip(hipMemcpyAsync(&w, ctx->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, st))) != SBP_OK) return fail(rc);
  if ((rc = from_hip(hipStreamSynchronize(st))) != SBP_OK) return fail(rc);
  if (w) {
    hipMemsetAsync(ctx->d_err, 0, sizeof(unsigned), st);
    return fail(status_from_errword(w));
  }
  *out = s.release();
  return SBP_OK;
}

int sbp_system_destroy(sbp_system s) {
  if (!s) return SBP_ERR_INVALID;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  destroy_graph(s);
  for (auto& h : s->blocks) {
    hipFree(h.d_metrics);
    hipFree(h.d_mpair);
    hipFree(h.d_x);
    hipFree(h.d_y);
    for (int r = 0; r < 4; ++r) hipFree(h.d_buf[r]);
    if (h.d_k) hipFree(h.d_k);
    if (h.d_rk) hipFree(h.d_rk);
  }
  if (s->d_blocks) hipFree(s->d_blocks);
  if (s->d_tiles) hipFree(s->d_tiles);
  if (s->d_pieces) hipFree(s->d_pieces);
  if (s->d_edata) hipFree(s->d_edata);
  if (s->d_evsegs) hipFree(s->d_evsegs);
  if (s->d_progs) hipFree(s->d_progs);
  if (s->d_time) hipFree(s->d_time);
  if (s->d_err) hipFree(s->d_err);
  if (s->d_partials) hipFree(s->d_partials);
  if (s->d_mixref) hipFree(s->d_mixref);
  if (s->d_tile_sums) hipFree(s->d_tile_sums);
  if (s->d_block_sums) hipFree(s->d_block_sums);
  if (s->h_block_sums) hipHostFree(s->h_block_sums);
  delete s;
  return SBP_OK;
}

int sbp_system_set_bc(sbp_system s, int block, int dir, int kind, int neighbour, const sbp_vortex_params* vortex) {
  if (!s || block < 0 || block >= (int)s->blocks.size() || dir < 0 || dir > 3) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  if (kind == SBP_BC_GRID) {
    if (neighbour < 0 || neighbour >= (int)s->blocks.size()) return SBP_ERR_INVALID;
    const HostBlock& o = s->blocks[neighbour];
    const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
    if ((ns ? h.nx : h.ny) != (ns ? o.nx : o.ny)) return SBP_ERR_EDGE;
  } else if (kind == SBP_BC_VORTEX) {
    if (!vortex || vortex->nvortices < 0 || vortex->nvortices > 5) return SBP_ERR_INVALID;
    h.bc_vortex[dir] = *vortex;
  } else if (kind != SBP_BC_THIS) {
    return SBP_ERR_INVALID;
  }
  h.bc_kind[dir] = kind;
  h.bc_nb[dir] = neighbour;
  h.bc_pieces[dir].clear();
  h.bc_prog[dir].reset();
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_bc_interpolate(sbp_system s, int block, int dir, int neighbour, int interp_op) {
  if (!s || block < 0 || block >= (int)s->blocks.size() || dir < 0 || dir > 3 || neighbour < 0 ||
      neighbour >= (int)s->blocks.size() || interp_op < 0 || interp_op >= INTERP_COUNT)
    return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  const HostBlock& o = s->blocks[neighbour];
  const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
  const int own = ns ? h.nx : h.ny, src = ns ? o.nx : o.ny;
  // fine to coarse iff this block has fewer points along ξ (lib.rs:818); the
  // operator's length assert then decides (interpolation4.rs:75,85)
  const bool f2c = h.nx < o.nx;
  if (!interp::coupling_valid(interp_op, f2c, f2c ? own : src, f2c ? src : own)) return SBP_ERR_EDGE;
  h.bc_kind[dir] = SBP_BC_INTERPOLATE;
  h.bc_nb[dir] = neighbour;
  h.bc_iop[dir] = interp_op;
  h.bc_pieces[dir].clear();
  h.bc_prog[dir].reset();
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_bc_multigrid(sbp_system s, int block, int dir, int npieces, const int* pieces) {
  if (!s || block < 0 || block >= (int)s->blocks.size() || dir < 0 || dir > 3 || npieces < 1 || !pieces)
    return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
  // lib.rs:829-849: piece k starts on the last point of piece k-1; the
  // pieces must end exactly on the last point of this edge
  int covered = 0;
  for (int k = 0; k < npieces; ++k) {
    const int g = pieces[3 * k], start = pieces[3 * k + 1], end = pieces[3 * k + 2];
    if (g < 0 || g >= (int)s->blocks.size()) return SBP_ERR_INVALID;
    const HostBlock& o = s->blocks[g];
    if (start < 0 || end > (ns ? o.nx : o.ny) || end - start < 1) return SBP_ERR_EDGE;
    covered += (end - start) - (k + 1 < npieces ? 1 : 0);
  }
  if (covered != (ns ? h.nx : h.ny)) return SBP_ERR_EDGE;
  h.bc_kind[dir] = SBP_BC_MULTIGRID;
  h.bc_prog[dir].reset();
  h.bc_pieces[dir].assign(pieces, pieces + 3 * (size_t)npieces);
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_bc_eval(sbp_system s, int block, int dir, sbp_expr e) {
  if (!s || !e || block < 0 || block >= (int)s->blocks.size() || dir < 0 || dir > 3) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  h.bc_kind[dir] = SBP_BC_EVAL;
  h.bc_prog[dir] = expr_blob(e);
  h.bc_pieces[dir].clear();
  s->dirty = true;
  return SBP_OK;
}

// A device copy of an expression program, freed with the returned guard.
struct DeviceProgram {
  void* p = nullptr;
  ~DeviceProgram() {
    if (p) hipFree(p);
  }
};

static int upload_program(sbp_expr e, hipStream_t st, DeviceProgram& out) {
  auto blob = expr_blob(e);
  if (!blob) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipMalloc(&out.p, blob->size()));
  SBP_TRY_HIP(hipMemcpyAsync(out.p, blob->data(), blob->size(), hipMemcpyHostToDevice, st));
  return SBP_OK;
}

int sbp_system_eval_block(sbp_system s, int block, double t, sbp_expr e) {
  if (!s || !e || block < 0 || block >= (int)s->blocks.size()) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  hipStream_t st = s->ctx->stream;
  DeviceProgram prog;
  SBP_TRY(upload_program(e, st, prog));
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(launch_eval_field(static_cast<const expr::Program*>(prog.p), t, h.d_x, h.d_y, h.d_buf[BUF_YN], h.ps, h.ld,
                                h.ny, h.nx, s->d_err, st));
  return check_err(s);
}

int sbp_system_error_eval(sbp_system s, sbp_expr e, double* out) {
  if (!s || !e || !out) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  hipStream_t st = s->ctx->stream;
  double t;
  SBP_TRY(sbp_system_get_time(s, &t));
  DeviceProgram prog;
  SBP_TRY(upload_program(e, st, prog));
  double err = 0.0;
  for (int i = 0; i < (int)s->blocks.size(); ++i) {
    HostBlock& h = s->blocks[i];
    // the accumulator buffer is free between steps: it holds the reference field
    double* ref = h.d_buf[BUF_ACC];
    SBP_TRY_HIP(launch_eval_field(static_cast<const expr::Program*>(prog.p), t * h.dt_scale, h.d_x, h.d_y, ref, h.ps, h.ld,
                                  h.ny, h.nx, s->d_err, st));
    long double sum;
    SBP_TRY(h2_sum(s, i, ref, nullptr, 0.0, sum));
    err += std::sqrt((double)sum);
  }
  SBP_TRY(check_err(s));
  *out = err;
  return SBP_OK;
}

int sbp_system_set_rhs_kind(sbp_system s, int kind) {
  if (!s || kind < SBP_RHS_AUTO || kind > SBP_RHS_UPWIND) return SBP_ERR_INVALID;
  for (auto& b : s->blocks) {
    int e;
    int saved = s->rhs_kind;
    s->rhs_kind = kind;
    block_upwind(s, b, &e);
    s->rhs_kind = saved;
    if (e) return e;
  }
  s->rhs_kind = kind;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_precision(sbp_system s, int precision) {
  if (!s || (precision != SBP_FP64 && precision != SBP_MIXED)) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  s->precision = precision;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_state(sbp_system s, int block, const double* host) {
  if (!s || !host || block < 0 || block >= (int)s->blocks.size()) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  for (int c = 0; c < 4; ++c)
    SBP_TRY_HIP(hipMemcpy2DAsync(h.d_buf[BUF_YN] + c * h.ps, h.ld * 8, host + (size_t)c * h.nx * h.ny, (size_t)h.nx * 8,
                                 (size_t)h.nx * 8, h.ny, hipMemcpyHostToDevice, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  return SBP_OK;
}

int sbp_system_get_state(sbp_system s, int block, double* host) {
  if (!s || !host || block < 0 || block >= (int)s->blocks.size()) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  for (int c = 0; c < 4; ++c)
    SBP_TRY_HIP(hipMemcpy2DAsync(host + (size_t)c * h.nx * h.ny, (size_t)h.nx * 8, h.d_buf[BUF_YN] + c * h.ps, h.ld * 8,
                                 (size_t)h.nx * 8, h.ny, hipMemcpyDeviceToHost, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  return SBP_OK;
}

int sbp_system_state_device(sbp_system s, int block, double** dev, int64_t* ld, int64_t* plane_stride) {
  if (!s || !dev || block < 0 || block >= (int)s->blocks.size()) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  *dev = h.d_buf[BUF_YN];
  if (ld) *ld = h.ld;
  if (plane_stride) *plane_stride = h.ps;
  return SBP_OK;
}

int sbp_system_get_metrics(sbp_system s, int block, double* host) {
  if (!s || !host || block < 0 || block >= (int)s->blocks.size()) return SBP_ERR_INVALID;
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  for (int c = 0; c < 5; ++c)
    SBP_TRY_HIP(hipMemcpy2DAsync(host + (size_t)c * h.nx * h.ny, (size_t)h.nx * 8, h.d_metrics + c * h.ps, h.ld * 8,
                                 (size_t)h.nx * 8, h.ny, hipMemcpyDeviceToHost, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  return SBP_OK;
}

int sbp_system_vortex(sbp_system s, double t, const sbp_vortex_params* params) {
  if (!s || !params || params->nvortices < 0 || params->nvortices > 5) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  s->parity = 0;  // every block's y_n is overwritten
  const VortexDev vp = to_dev(*params);
  for (auto& h : s->blocks)
    SBP_TRY_HIP(launch_vortex(vp, gamma, t, h.d_x, h.d_y, h.d_buf[BUF_YN], h.ps, h.ld, h.ny, h.nx, s->d_err, s->ctx->stream));
  return check_err(s);
}

int sbp_system_vortex_block(sbp_system s, int block, double t, const sbp_vortex_params* params) {
  if (!s || !params || block < 0 || block >= (int)s->blocks.size() || params->nvortices < 0 || params->nvortices > 5)
    return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(normalize(s));
  const VortexDev vp = to_dev(*params);
  HostBlock& h = s->blocks[block];
  SBP_TRY_HIP(launch_vortex(vp, gamma, t, h.d_x, h.d_y, h.d_buf[BUF_YN], h.ps, h.ld, h.ny, h.nx, s->d_err, s->ctx->stream));
  return check_err(s);
}

int sbp_system_rhs(sbp_system s, double t, double* k_host) {
  if (!s || !k_host) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  bool alloc = false;
  for (auto& h : s->blocks)
    if (!h.d_k) {
      SBP_TRY_HIP(hipMalloc(&h.d_k, sizeof(double) * 4 * h.ps));
      SBP_TRY_HIP(hipMemsetAsync(h.d_k, 0, sizeof(double) * 4 * h.ps, st));
      alloc = true;
    }
  if (alloc) s->dirty = true;
  SBP_TRY(normalize(s));
  SBP_TRY(rebuild(s));
  SBP_TRY_HIP(hipMemcpyAsync(s->d_time_rhs, &t, sizeof(double), hipMemcpyHostToDevice, st));
  StageArgs a = stage_args(s, 0.0, 0, 0);
  a.time = s->d_time_rhs;
  a.rk_b = a.rk_a = 0.0;
  for (const Group& g : s->groups) SBP_TRY_HIP(launch_group(s, g, g.shape, g.sched, MODE_RHS, a, st));
  size_t off = 0;
  for (auto& h : s->blocks) {
    for (int c = 0; c < 4; ++c)
      SBP_TRY_HIP(hipMemcpy2DAsync(k_host + off + (size_t)c * h.nx * h.ny, (size_t)h.nx * 8, h.d_k + c * h.ps, h.ld * 8,
                                   (size_t)h.nx * 8, h.ny, hipMemcpyDeviceToHost, st));
    off += (size_t)4 * h.nx * h.ny;
  }
  return check_err(s);
}

int sbp_system_advance(sbp_system s, double dt, int64_t nsteps) {
  if (!s || nsteps < 0 || !std::isfinite(dt)) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(advance_impl(s, dt, nsteps));
  return check_err(s);
}

int sbp_system_get_time(sbp_system s, double* t) {
  if (!s || !t) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipMemcpyAsync(t, s->d_time, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  return SBP_OK;
}

int sbp_system_set_time(sbp_system s, double t) {
  if (!s) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipMemcpyAsync(s->d_time, &t, sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  SBP_TRY_HIP(hipStreamSynchronize(s->ctx->stream));
  return SBP_OK;
}

int sbp_system_max_dt(sbp_system s, double* out) {
  if (!s || !out) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  bool is_h2 = false;
  for (auto& h : s->blocks) {
    OpInfo oe, ox;
    op_info(h.op_eta, oe);
    op_info(h.op_xi, ox);
    is_h2 = is_h2 || oe.h2 || ox.h2;
  }
  const double c_max = is_h2 ? 0.5 : 1.0;
  double max_dt = INFINITY;
  SBP_TRY(normalize(s));
  hipStream_t st = s->ctx->stream;
  SBP_TRY(ensure_partials(s));
  for (auto& h : s->blocks) {
    SBP_TRY_HIP(launch_maxspeed(h.d_buf[BUF_YN], h.d_metrics, h.ps, h.ld, h.ny, h.nx, s->d_partials, st));
    double r[2];
    SBP_TRY(reduce_result(s->d_partials, st, r));
    const double mu = r[0], mv = r[1];
    const double dx = 1.0 / (double)h.nx, dy = 1.0 / (double)h.ny;
    const double c_dt = std::fmax(mu / dx, mv / dy);
    max_dt = std::fmin(max_dt, c_max / c_dt);
  }
  *out = max_dt;
  return SBP_OK;
}

int sbp_system_error_vortex(sbp_system s, const sbp_vortex_params* params, double* out) {
  if (!s || !params || !out || params->nvortices < 0 || params->nvortices > 5) return SBP_ERR_INVALID;
  if (double g; !gamma_get(g)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  double t;
  SBP_TRY(sbp_system_get_time(s, &t));
  const VortexDev vp = to_dev(*params);
  double e = 0.0;
  for (int i = 0; i < (int)s->blocks.size(); ++i) {
    long double sum;
    SBP_TRY(h2_sum(s, i, nullptr, &vp, t * s->blocks[i].dt_scale, sum));
    e += std::sqrt((double)sum);
  }
  *out = e;
  return SBP_OK;
}

int sbp_system_error_vortex_block(sbp_system s, int block, const sbp_vortex_params* params, double* out) {
  if (!s || !params || !out || block < 0 || block >= (int)s->blocks.size() || params->nvortices < 0 ||
      params->nvortices > 5)
    return SBP_ERR_INVALID;
  if (double g; !gamma_get(g)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  double t;
  SBP_TRY(sbp_system_get_time(s, &t));
  const VortexDev vp = to_dev(*params);
  long double sum;
  SBP_TRY(h2_sum(s, block, nullptr, &vp, t * s->blocks[block].dt_scale, sum));
  *out = std::sqrt((double)sum);
  return SBP_OK;
}

int sbp_system_set_dt_scale(sbp_system s, int block, double scale) {
  if (!s || block < 0 || block >= (int)s->blocks.size() || !(scale > 0.0) || !std::isfinite(scale)) return SBP_ERR_INVALID;
  s->blocks[block].dt_scale = scale;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_energy(sbp_system s, double* out) {
  if (!s || !out) return SBP_ERR_INVALID;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  long double e = 0.0L;
  for (int i = 0; i < (int)s->blocks.size(); ++i) {
    long double sum;
    SBP_TRY(h2_sum(s, i, nullptr, nullptr, 0.0, sum));
    e += sum;
  }
  *out = (double)e;
  return SBP_OK;
}

int sbp_system_check(sbp_system s) {
  if (!s) return SBP_ERR_INVALID;
  return check_err(s);
}

int sbp_system_profile(sbp_system s, double dt, int64_t nsteps, double* ms_total, double* ms_stage_avg) {
  if (!s || nsteps <= 0 || !ms_total || !ms_stage_avg) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(rebuild(s));
  hipStream_t st = s->ctx->stream;
  hipEvent_t e0, e1;
  SBP_TRY_HIP(hipEventCreate(&e0));
  SBP_TRY_HIP(hipEventCreate(&e1));
  // graphs are captured before the timed region
  const int G = steps_per_graph(s);
  if (nsteps >= G) SBP_TRY(ensure_graph(s, dt, G));
  SBP_TRY_HIP(hipStreamSynchronize(st));
  SBP_TRY_HIP(hipEventRecord(e0, st));
  SBP_TRY(advance_impl(s, dt, nsteps));
  SBP_TRY_HIP(hipEventRecord(e1, st));
  SBP_TRY_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SBP_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_total = ms;
  // per-launch durations of the stage kernels: up to 8 more steps with every
  // stage launch bracketed by its own events
  std::vector<hipEvent_t> ev;
  double sum = 0.0;
  int nl = 0;
  const int64_t tsteps = std::min<int64_t>(nsteps, 8);
  for (int64_t step = 0; step < tsteps; ++step) {
    for (int stage = 0; stage < 4; ++stage) {
      const StageArgs a = stage_args(s, dt, stage, s->parity);
      for (const Group& g : s->groups) {
        hipEvent_t a0, a1;
        SBP_TRY_HIP(hipEventCreate(&a0));
        SBP_TRY_HIP(hipEventCreate(&a1));
        SBP_TRY_HIP(hipEventRecord(a0, st));
        SBP_TRY_HIP(launch_group(s, g, g.shape, g.sched, stage, a, st));
        SBP_TRY_HIP(hipEventRecord(a1, st));
        ev.push_back(a0);
        ev.push_back(a1);
      }
    }
    SBP_TRY_HIP(launch_advance_time(s->d_time, dt, st));
    if (s->precision == SBP_MIXED) s->parity ^= 1;
  }
  SBP_TRY_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < ev.size(); i += 2) {
    float m = 0.f;
    SBP_TRY_HIP(hipEventElapsedTime(&m, ev[i], ev[i + 1]));
    sum += m;
    ++nl;
    hipEventDestroy(ev[i]);
    hipEventDestroy(ev[i + 1]);
  }
  *ms_stage_avg = nl ? sum / nl : 0.0;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return check_err(s);
}

int sbp_system_advance_adaptive(sbp_system s, double dt, double* guess_dt, double maxerr) {
  if (!s || !guess_dt || !std::isfinite(dt) || !std::isfinite(*guess_dt) || !(*guess_dt > 0.0) || std::isnan(maxerr))
    return SBP_ERR_INVALID;
  if (s->precision != SBP_FP64) return SBP_ERR_INVALID;
  // one trial step for the whole system: per-block step scales (batched
  // independent solves) have no meaning here
  for (auto& h : s->blocks)
    if (h.dt_scale != 1.0) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  SBP_TRY(normalize(s));
  SBP_TRY(rebuild(s));
  const int nb = (int)s->blocks.size();
  if (s->tile_sums_cap < s->total_tiles) {
    if (s->d_tile_sums) hipFree(s->d_tile_sums);
    s->d_tile_sums = nullptr;
    s->tile_sums_cap = 0;
    SBP_TRY_HIP(hipMalloc(&s->d_tile_sums, sizeof(double) * 2 * s->total_tiles));
    s->tile_sums_cap = s->total_tiles;
  }
  if (!s->d_block_sums) SBP_TRY_HIP(hipMalloc(&s->d_block_sums, sizeof(double) * 2 * nb));
  if (!s->h_block_sums) SBP_TRY_HIP(hipHostMalloc(&s->h_block_sums, sizeof(double) * 2 * nb, 0));
  // Each trial step is four fused stage launches per group (modes BS0-BS3,
  // the stage kernel's finish), then one workgroup per block sums its tiles'
  // error terms. BogackiShampine (integrate/src/lib.rs:138-151): C carries
  // its leading 0, so integrate's stage time t + g·C[i-1] (:190-196) places
  // stage i at the previous stage's node; the next-stage weights 1/2, 3/4
  // and the solution's 4/9 (with the tableau identities, see the kernel),
  // the embedded solution's 1/8 of the last derivative.
  static const double C[4] = {0.0, 1.0 / 2.0, 3.0 / 4.0, 1.0};
  static const double Anext[3] = {1.0 / 2.0, 3.0 / 4.0, 4.0 / 9.0};
  // advance_adaptive (euler/src/lib.rs:149-185): the loop's time starts at 0
  // and grows by the trial step inside integrate and once more on acceptance;
  // the error is Field::h2_err of the previous state against the embedded
  // (second) solution, summed over blocks (not in quadrature, as
  // System::error); accepted steps grow the guess by 1.05, rejected ones
  // shrink it by 0.8. The boundary data sees the loop's stage times. An
  // accepted solution lives in the other FP64 buffer (`swap`) and is copied
  // back once at the end.
  double time = 0.0, g = *guess_dt;
  int swap = 0;
  while (time < dt) {
    for (int i = 0; i < 4; ++i) {
      StageArgs a = stage_args(s, g, 0, 0);
      a.time = nullptr;
      a.time_offset = i == 0 ? time : time + g * C[i - 1];
      a.rk_a = i < 3 ? Anext[i] : 0.0;
      a.rk_b = i == 3 ? 1.0 / 8.0 : 0.0;
      a.swap = swap;
      a.partials = s->d_tile_sums;
      for (const Group& grp : s->groups) SBP_TRY_HIP(launch_group(s, grp, grp.shape, grp.sched, MODE_BS0 + i, a, st));
    }
    SBP_TRY_HIP(launch_block_sums(s->d_blocks, nb, s->d_tile_sums, s->d_block_sums, st));
    SBP_TRY_HIP(hipMemcpyAsync(s->h_block_sums, s->d_block_sums, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st));
    SBP_TRY_HIP(hipStreamSynchronize(st));
    double err = 0.0;
    for (int bi = 0; bi < nb; ++bi) {
      const long double sum = (long double)s->h_block_sums[2 * bi] + (long double)s->h_block_sums[2 * bi + 1];
      err += std::sqrt((double)sum);
    }
    time += g;
    if (err < maxerr) {
      time += g;
      swap ^= 1;
      g *= 1.05;
    } else {
      g *= 0.8;
    }
  }
  if (swap)
    for (auto& h : s->blocks)
      SBP_TRY_HIP(launch_copy_planes(h.d_buf[BUF_ACC], h.ps, h.ld, h.d_buf[BUF_YN], h.ps, h.ld, 4, h.ny, h.nx, st));
  *guess_dt = g;
  SBP_TRY_HIP(hipStreamSynchronize(st));
  return check_err(s);
}

int sbp_system_integrate(sbp_system s, int nstages, const double* A, const double* B, const double* C, double dt,
                         int64_t nsteps) {
  if (!s || !A || !B || (nstages > 1 && !C) || nstages < 1 || nstages + 1 > kMaxWorkSets || nsteps < 0 ||
      !std::isfinite(dt))
    return SBP_ERR_INVALID;
  if (s->precision != SBP_FP64) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  SBP_TRY(ensure_work_sets(s, nstages + 1));
  SBP_TRY(normalize(s));
  SBP_TRY(rebuild(s));
  const int SET_STAGE = nstages;  // the stage input after k0..k{s-1}
  auto set_of = [](const HostBlock& h, int r) { return h.d_rk + (int64_t)r * 4 * h.ps; };
  // out = prev + Σ (w_j·dt_b)·k_j over the nonzero weights, in order (the
  // reference's clone_from and scaled_adds, integrate/src/lib.rs:187-211)
  auto combine = [&](const double* w, int nw, double* const* out) -> int {
    for (size_t bi = 0; bi < s->blocks.size(); ++bi) {
      const HostBlock& h = s->blocks[bi];
      const double g = dt * h.dt_scale;
      LinComb lc{};
      for (int j = 0; j < nw; ++j) {
        if (w[j] == 0.0) continue;
        lc.k[lc.nk] = set_of(h, j);
        lc.s[lc.nk] = w[j] * g;
        ++lc.nk;
      }
      SBP_TRY_HIP(launch_lincomb(out[bi], h.d_buf[BUF_YN], lc, h.ps, h.ld, h.ny, h.nx, st));
    }
    return SBP_OK;
  };
  std::vector<double*> stage_out(s->blocks.size()), final_out(s->blocks.size());
  for (size_t bi = 0; bi < s->blocks.size(); ++bi) {
    stage_out[bi] = set_of(s->blocks[bi], SET_STAGE);
    final_out[bi] = s->blocks[bi].d_buf[BUF_YN];
  }
  // integrate (utils/integrate/src/lib.rs:176-222): stage 0 at t from y_n,
  // stage i > 0 at t + dt·C[i-1] from y_n + Σ_{j<i} (A[i][j]·dt)·k_j, then
  // y_{n+1} = y_n + Σ (B[j]·dt)·k_j and t += dt (the device time, read by the
  // far-field edge data)
  for (int64_t step = 0; step < nsteps; ++step) {
    for (int i = 0; i < nstages; ++i) {
      if (i > 0) SBP_TRY(combine(A + (size_t)i * nstages, i, stage_out.data()));
      StageArgs a = stage_args(s, 0.0, 0, 0);
      a.time_offset = i == 0 ? 0.0 : dt * C[i - 1];
      a.rk_a = a.rk_b = 0.0;
      a.rhs_in_sel = i == 0 ? 0 : SET_STAGE + 1;
      a.rhs_out_sel = i + 1;
      for (const Group& grp : s->groups) SBP_TRY_HIP(launch_group(s, grp, grp.shape, grp.sched, MODE_RHS, a, st));
    }
    SBP_TRY(combine(B, nstages, final_out.data()));
    SBP_TRY_HIP(launch_advance_time(s->d_time, dt, st));
  }
  return check_err(s);
}

int sbp_system_advance_timed(sbp_system s, double dt, int64_t nsteps, double* ms) {
  if (!s || nsteps <= 0 || !ms || !std::isfinite(dt)) return SBP_ERR_INVALID;
  double gamma;
  if (!gamma_get(gamma)) return SBP_ERR_GAMMA;
  SBP_TRY_HIP(hipSetDevice(s->ctx->device));
  SBP_TRY(rebuild(s));
  hipStream_t st = s->ctx->stream;
  const int G = steps_per_graph(s);
  if (nsteps >= G) SBP_TRY(ensure_graph(s, dt, G));
  hipEvent_t e0, e1;
  SBP_TRY_HIP(hipEventCreate(&e0));
  SBP_TRY_HIP(hipEventCreate(&e1));
  hipError_t he = hipEventRecord(e0, st);
  int rc = he == hipSuccess ? advance_impl(s, dt, nsteps) : from_hip(he);
  if (rc == SBP_OK) he = hipEventRecord(e1, st);
  if (rc == SBP_OK && he == hipSuccess) he = hipEventSynchronize(e1);
  float m = 0.f;
  if (rc == SBP_OK && he == hipSuccess) he = hipEventElapsedTime(&m, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != SBP_OK) return rc;
  if (he != hipSuccess) return from_hip(he);
  *ms = m;
  return check_err(s);
}

int sbp_system_tile_shape(sbp_system s, int* tx, int* ty, int* ry) {
  if (!s) return SBP_ERR_INVALID;
  SBP_TRY(rebuild(s));
  const TileShape& t = kTileShapes[s->groups.empty() ? 0 : s->groups[0].shape];
  if (tx) *tx = t.tx;
  if (ty) *ty = t.ty;
  if (ry) *ry = t.ry;
  return SBP_OK;
}

int sbp_system_tile_index(sbp_system s, int* index) {
  if (!s || !index) return SBP_ERR_INVALID;
  SBP_TRY(rebuild(s));
  *index = s->groups.empty() ? 0 : s->groups[0].shape;
  return SBP_OK;
}

int sbp_system_set_tile_shape(sbp_system s, int shape_index) {
  if (!s || shape_index < -1 || shape_index >= kNumTileShapes) return SBP_ERR_INVALID;
  s->forced_shape = shape_index;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_set_schedule(sbp_system s, int schedule) {
  if (!s || schedule < -1 || schedule > 2) return SBP_ERR_INVALID;
  s->forced_schedule = schedule;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_schedule(sbp_system s, int* schedule) {
  if (!s || !schedule) return SBP_ERR_INVALID;
  SBP_TRY(rebuild(s));
  *schedule = s->groups.empty() ? 0 : s->groups[0].sched;
  return SBP_OK;
}

int sbp_system_set_tile_order(sbp_system s, int order) {
  if (!s || order < -1 || order > 1) return SBP_ERR_INVALID;
  s->forced_order = order;
  s->dirty = true;
  return SBP_OK;
}

int sbp_system_tile_order(sbp_system s, int* order) {
  if (!s || !order) return SBP_ERR_INVALID;
  SBP_TRY(rebuild(s));
  *order = s->groups.empty() ? 0 : s->groups[0].order;
  return SBP_OK;
}

int sbp_tile_shape_count(void) { return kNumTileShapes; }

int sbp_tile_shape_get(int index, int* tx, int* ty, int* ry) {
  if (index < 0 || index >= kNumTileShapes) return SBP_ERR_INVALID;
  if (tx) *tx = kTileShapes[index].tx;
  if (ty) *ty = kTileShapes[index].ty;
  if (ry) *ry = kTileShapes[index].ry;
  return SBP_OK;
}

}  // extern "C"
