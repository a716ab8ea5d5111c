// Device-side descriptors shared by the HIP kernels and the host runtime.
#pragma once
#include <stdint.h>

#include "../../include/sbpgpu.h"

namespace sbp {
namespace expr {
struct Program;
}

// Work sets of the Runge-Kutta paths that keep their stage derivatives.
constexpr int kMaxWorkSets = 10;

// Roles of the four state buffers of a block during one fused RK4 step.
// FP64 form: y_n and the stage inputs Y2 (YA), Y3 (YB) and Y4 (YA, written
// by stage 3 over Y2); stage 2 also writes W = Y2 + 2 Y3 - y_n into ACC,
// from which and Y4 the last stage forms y_{n+1} (see the stage kernel's
// finish). Y2, Y3, Y4 and W are stored point-major (the four components of a
// point adjacent), y_n in planes.
// Mixed form: YA holds the FP32 stage increments Δ_1 = Y2 - y_n and
// Δ_2 = Y3 - y_n (two sets of 4·ps floats, the second at float offset 4·ps;
// the four components of a point are adjacent), YB holds Δ_3 = Y4 - y_n in
// its first half; the last stage writes y_{n+1} into the ACC buffer, after
// which y_n and ACC swap roles (StageArgs::swap).
enum BufRole : int { BUF_YN = 0, BUF_ACC = 1, BUF_YA = 2, BUF_YB = 3 };

// Boundary-data source of one block edge inside a kernel.
enum EdgeSrc : int {
  EDGE_THIS = SBP_BC_THIS,      // own opposite edge of the stage input (periodic)
  EDGE_GRID = SBP_BC_GRID,      // neighbour block's opposite edge of the stage input
  EDGE_VORTEX = SBP_BC_VORTEX,  // analytic vortex on this edge's coordinates at stage time
  EDGE_INTERP = SBP_BC_INTERPOLATE,  // neighbour's opposite edge through a 2:1 interpolation operator
  EDGE_MULTIGRID = SBP_BC_MULTIGRID,  // pieces of several neighbours' opposite edges
  EDGE_EVAL = SBP_BC_EVAL,      // expressions of (t, x, y) on this edge's coordinates at stage time
  EDGE_EXPLICIT = 100,          // caller-provided 4 x n array (operator-level rhs)
  EDGE_NONE = 101,              // no SAT on this edge (RHS_no_SAT)
};

// Stage-kernel mode: RHS only, one of the four fused RK4 stages, or one of
// the four fused stages of an embedded Bogacki-Shampine trial step
// (advance_adaptive; FP64 only): BS0 writes Y2 into YA, BS1 Y3 into YB, BS2
// the solution into the FP64 buffer that does not hold y_n (StageArgs::swap
// says which does), BS3 writes no state, only the H-norm sums of
// (y_n - embedded solution) of each tile into StageArgs::partials.
enum StageMode : int {
  MODE_RHS = -1,
  MODE_S0 = 0,
  MODE_S1 = 1,
  MODE_S2 = 2,
  MODE_S3 = 3,
  MODE_BS0 = 4,
  MODE_BS1 = 5,
  MODE_BS2 = 6,
  MODE_BS3 = 7
};

struct VortexDev {
  int32_t n;
  int32_t pad_;
  double x0[5], y0[5], rstar[5], eps[5];
  double mach;
};

// One block (grid patch) as the kernels see it. All planes share a row pitch
// `ld` (elements) and plane stride `ps`; state buffers hold 4 planes
// (rho, rhou, rhov, e), metrics 5 planes (J, J·ξx, J·ξy, J·ηx, J·ηy).
struct BlockDev {
  int nx, ny, ld, op_xi, op_eta, upwind;
  int ntx, nty;  // tiles along ξ and η for the active tile shape
  int64_t tile_off;  // slot of the block's first tile in per-tile arrays (StageArgs::partials, mixref)
  int64_t ps;
  const double* metrics;
  // metric pairs as the stage kernel reads them: (J·ξx, J·ξy) of point o at
  // doubles 2o, 2o + 1, (J·ηx, J·ηy) at 2ps + 2o, 2ps + 2o + 1
  const double* mpair;
  const double* xc;  // ξ/η grid coordinates x(η, ξ), y(η, ξ)
  const double* yc;
  double* buf[4];  // BufRole
  double* rhs_in;  // MODE_RHS input state (4 planes)
  double* rhs_out;  // MODE_RHS output k (4 planes)
  // Runge-Kutta work sets (4 planes each): advance_adaptive uses k0..k3, its
  // stage input (set 4) and second solution (set 5); a tableau of s stages
  // (sbp_system_integrate) uses k0..k{s-1} and its stage input (set s)
  double* rk[kMaxWorkSets];
  double idx_xi, idx_eta;  // 1/dξ and 1/dη (algos.rs:125-130)
  double hi_xi, hi_eta;    // SAT factors (n-1)/H0 (lib.rs:1034-1038,1095-1099)
  double dt_scale;         // this block steps with dt·dt_scale; its time is t·dt_scale
  int edge_src[4];         // EdgeSrc per sbp_direction
  int edge_nb[4];          // neighbour block for EDGE_GRID
  const double* edge_ptr[4];  // EDGE_EXPLICIT data
  int64_t edge_cs[4];         // component stride of EDGE_EXPLICIT data
  VortexDev vortex[4];        // EDGE_VORTEX parameters
  const expr::Program* eprog[4];  // EDGE_EVAL program (device copy)
  double* edata[4];               // EDGE_EVAL data at the current stage time (4 x edge length)
  // characteristic SAT of each edge at the current stage (4 x edge length,
  // component-major) when the edge-SAT kernel evaluates them
  // (StageArgs::sat_buf)
  double* sat[4];
  int edge_iop[4];            // EDGE_INTERP: interpolation operator (InterpId)
  int edge_f2c[4];            // EDGE_INTERP: 1 = fine to coarse
  const int* edge_pieces[4];  // EDGE_MULTIGRID: (block, start, end) triples
  int edge_npieces[4];
};

// One block edge in a per-stage edge list (the edge-SAT kernel, the Eval
// edges' data kernel).
struct SatSeg {
  int32_t block, dir;
};

struct TileDev {
  int32_t block;
  int16_t tx, ty;
};

struct StageArgs {
  const BlockDev* blocks;
  const TileDev* tiles;
  const double* time;  // device-resident simulation time (integrate's `time`)
  unsigned int* error;  // sticky error word (bit per sbp_status)
  double dt;
  double gamma;
  double rk_b;  // b_s: the stage's weight in the accumulator (b_s·dt_block)
  double rk_a;  // a_s: weight of the next stage input (a_s·dt_block)
  double time_offset;  // c_s·dt added to *time for boundary data
  int mixed;  // FP32 right-hand side with FP64 state (SBP_MIXED)
  int swap;   // mixed form (odd steps) and the BS modes: y_n lives in BUF_ACC, the other FP64 buffer is BUF_YN
  int sat_buf;  // 1: the edge tiles read the edge SATs from BlockDev::sat (edge-SAT kernel)
  double* partials;  // MODE_BS3: (sum, compensation) per tile, at BlockDev::tile_off + ty·ntx + tx
  int rhs_in_sel, rhs_out_sel;  // MODE_RHS buffers: 0 = rhs_in / rhs_out, s > 0 = rk[s - 1]
  // mixed form: reference face terms of every tile (16 doubles at slot
  // BlockDev::tile_off + ty·ntx + tx), written by mix_ref_kernel before the
  // first stage of a step (and before a right-hand-side launch)
  double* mixref;
  // RK4 step without a separate time update: thread 0 of workgroup 0 of one
  // dispatch writes *time_w = *time (stage S0: the step's start time into the
  // slot stages S1-S3 read) or *time + dt (stage S1: the next step's time into
  // the slot S0 reads); no dispatch of the same stage reads *time_w. Null:
  // no write.
  double* time_w;
};

}  // namespace sbp
