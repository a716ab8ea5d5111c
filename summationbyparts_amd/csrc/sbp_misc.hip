// Operator application, metrics, vortex evaluation, reductions and small
// helpers (HIP, gfx950).
#include <hip/hip_runtime.h>

#include "sbp_common.hpp"
#include "sbp_kernels.hpp"

namespace sbp {

const TileShape kTileShapes[kNumTileShapes] = {
    {32, 16, 1, 2, 1}, {32, 16, 1, 2, 0}, {32, 24, 1, 2, 0}, {32, 20, 1, 2, 0}, {32, 12, 1, 2, 0},
    {64, 16, 1, 2, 0}, {32, 24, 1, 2, 1}, {16, 16, 1, 2, 0}, {16, 24, 1, 2, 0}, {32, 8, 1, 1, 0},
    {32, 16, 1, 1, 0}, {64, 8, 1, 1, 0},  {32, 12, 1, 1, 0},
};

// ---------------------------------------------------------------------------
// Operator application along one axis (SbpOperator2d::diffxi/diffeta,
// UpwindOperator2d::dissxi/disseta; algos.rs:101-187 and :371-519). Each
// thread computes kApplyR consecutive outputs along the axis from one window
// of kApplyR + D - 1 inputs (closure rows, the first and last M, take the
// start or end block per output). Along ξ a workgroup first stages its four
// rows (256 outputs each, with N columns of margin on both sides, enough for
// any closure row) and the start block in LDS, and its outputs in LDS for
// coalesced stores; along η
// the window is a column, its loads coalesced across the wavefront's
// columns.
constexpr int kApplyR = 4;

template <class B>
__device__ __forceinline__ void window_sums(const double* w, int stride, double out[kApplyR]) {
#pragma unroll
  for (int r = 0; r < kApplyR; ++r) out[r] = 0.0;
#pragma unroll
  for (int k = 0; k < kApplyR + B::D - 1; ++k) {
    const double v = w[k * stride];
#pragma unroll
    for (int r = 0; r < kApplyR; ++r) {
      const int d = k - r;
      if (d >= 0 && d < B::D && B::diag(d) != 0.0) out[r] += B::diag(d) * v;
    }
  }
}

// LDS index of row element k along ξ: one pad double after every four, so
// that lanes four elements apart (each thread's window starts four columns
// after its neighbour's) hit different banks
__device__ __forceinline__ constexpr int xi_pad(int k) { return k + (k >> 2); }

template <class OP, int KIND>
__global__ __launch_bounds__(256) void apply_xi_kernel(const double* __restrict__ prev, int64_t ldp, double* __restrict__ fut,
                                                       int64_t ldf, int ny, int nx, double idx) {
  using B = Blk<OP, KIND>;
  // margin HM = N columns on both sides: an end-closure row reads up to N - 1
  // columns to its left
  constexpr int TW = 64 * kApplyR, HM = B::N, LW = TW + 2 * HM, C = HM - B::H;
  __shared__ double rows[4][xi_pad(LW) + 1];
  __shared__ double outs[4][xi_pad(TW)];
  __shared__ double tab[B::M * B::N];  // the start block, for closure rows
  const int lane = threadIdx.x & 63, ry = threadIdx.x >> 6;
  const int x0 = blockIdx.x * TW, y = blockIdx.y * 4 + ry;
  const int r0 = x0 - HM;  // LDS element k holds column r0 + k
  const bool clos = x0 < B::M || x0 + TW > nx - B::M;  // the workgroup writes closure columns
  if (clos)
    for (int e = threadIdx.x; e < B::M * B::N; e += 256) tab[e] = B::start(e / B::N, e % B::N);
  if (y < ny) {
    const double* p = prev + (int64_t)y * ldp;
    for (int k = lane; k < LW; k += 64) {
      const int gx = r0 + k;
      rows[ry][xi_pad(k)] = (gx >= 0 && gx < nx) ? p[gx] : 0.0;
    }
  }
  __syncthreads();
  // outputs go through LDS too, so that the stores are coalesced
  const int xa = x0 + lane * kApplyR;
  const double* l = rows[ry];
  if (y < ny && xa < nx) {
    double o[kApplyR];
    if (xa >= B::M && xa + kApplyR <= nx - B::M) {
      // window: elements 4·lane + C + k, k < kApplyR + D - 1
      const double* w = l + 5 * lane;
#pragma unroll
      for (int r = 0; r < kApplyR; ++r) o[r] = 0.0;
#pragma unroll
      for (int k = 0; k < kApplyR + B::D - 1; ++k) {
        const double v = w[xi_pad(C + k)];
#pragma unroll
        for (int r = 0; r < kApplyR; ++r) {
          const int d = k - r;
          if (d >= 0 && d < B::D && B::diag(d) != 0.0) o[r] += B::diag(d) * v;
        }
      }
    } else {
      // closure columns (start or end block, end rows read the start block
      // flipped with the kind's sign), interior columns of the same thread,
      // and columns past the end
#pragma unroll
      for (int r = 0; r < kApplyR; ++r) {
        const int i = xa + r;
        double sum = 0.0;
        if (i < B::M) {
#pragma unroll
          for (int k = 0; k < B::N; ++k) sum += tab[i * B::N + k] * l[xi_pad(k - r0)];
        } else if (i >= nx - B::M && i < nx) {
          const int rr = B::M - 1 - (i - (nx - B::M));
#pragma unroll
          for (int k = 0; k < B::N; ++k)
            sum += (B::end_sign * tab[rr * B::N + (B::N - 1 - k)]) * l[xi_pad(nx - B::N + k - r0)];
        } else if (i < nx) {
#pragma unroll
          for (int d = 0; d < B::D; ++d)
            if (B::diag(d) != 0.0) sum += B::diag(d) * l[xi_pad(i - B::H + d - r0)];
        }
        o[r] = sum;
      }
    }
#pragma unroll
    for (int r = 0; r < kApplyR; ++r) outs[ry][xi_pad(lane * kApplyR + r)] = o[r] * idx;
  }
  __syncthreads();
  if (y >= ny) return;
  double* f = fut + (int64_t)y * ldf;
#pragma unroll
  for (int r = 0; r < kApplyR; ++r) {
    const int x = x0 + r * 64 + lane;
    if (x < nx) f[x] = outs[ry][xi_pad(r * 64 + lane)];
  }
}

template <class OP, int KIND>
__global__ __launch_bounds__(256) void apply_eta_kernel(const double* __restrict__ prev, int64_t ldp, double* __restrict__ fut,
                                                        int64_t ldf, int ny, int nx, double idx) {
  using B = Blk<OP, KIND>;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int ya = (blockIdx.y * 4 + (threadIdx.x >> 6)) * kApplyR;
  if (x >= nx || ya >= ny) return;
  const double* p = prev + x;
  double o[kApplyR];
  if (ya >= B::M && ya + kApplyR <= ny - B::M) {
    window_sums<B>(p + (int64_t)(ya - B::H) * ldp, (int)ldp, o);
  } else {
#pragma unroll
    for (int r = 0; r < kApplyR; ++r) o[r] = ya + r < ny ? row_sum<B>(p, (int)ldp, ya + r, ny, 0) : 0.0;
  }
#pragma unroll
  for (int r = 0; r < kApplyR; ++r)
    if (ya + r < ny) fut[(int64_t)(ya + r) * ldf + x] = o[r] * idx;
}

template <class OP, int KIND>
static void launch_apply_axis(int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny, int nx, double idx,
                              hipStream_t stream) {
  if (axis == 0) {
    dim3 grid((nx + 64 * kApplyR - 1) / (64 * kApplyR), (ny + 3) / 4);
    apply_xi_kernel<OP, KIND><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
  } else {
    dim3 grid((nx + 63) / 64, (ny + 4 * kApplyR - 1) / (4 * kApplyR));
    apply_eta_kernel<OP, KIND><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
  }
}

template <class OP>
static hipError_t apply_op_t(int kind, int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny, int nx,
                             hipStream_t stream) {
  const int n = axis == 0 ? nx : ny;
  const double dx = OP::is_h2 ? 1.0 / (double)(n - 2) : 1.0 / (double)(n - 1);
  const double idx = 1.0 / dx;
  if (kind == SBP_DIFF) {
    launch_apply_axis<OP, SBP_DIFF>(axis, prev, ldp, fut, ldf, ny, nx, idx, stream);
  } else {
    if constexpr (OP::is_upwind) {
      launch_apply_axis<OP, SBP_DISS>(axis, prev, ldp, fut, ldf, ny, nx, idx, stream);
    } else {
      return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_apply_op(int op, int kind, int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny,
                           int nx, hipStream_t stream) {
  switch (op) {
    case UPWIND4: return apply_op_t<ops::Upwind4>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    case UPWIND4H2: return apply_op_t<ops::Upwind4h2>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    case UPWIND9: return apply_op_t<ops::Upwind9>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    case UPWIND9H2: return apply_op_t<ops::Upwind9h2>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    case SBP4: return apply_op_t<ops::SBP4>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    case SBP8: return apply_op_t<ops::SBP8>(kind, axis, prev, ldp, fut, ldf, ny, nx, stream);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Metrics from coordinate derivatives (grid.rs:110-131).
__global__ __launch_bounds__(256) void metrics_finish_kernel(const double* __restrict__ x_xi, const double* __restrict__ x_eta,
                                                             const double* __restrict__ y_xi, const double* __restrict__ y_eta,
                                                             double* __restrict__ met, int64_t ps, int64_t ld, int ny, int nx,
                                                             unsigned* err) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  const int64_t o = (int64_t)y * ld + x;
  const double a = x_xi[o], b = x_eta[o], c = y_xi[o], d = y_eta[o];
  const double detj = a * d - b * c;
  if (!(detj > 0.0)) atomicOr(err, ERRBIT_DETJ);
  met[o] = detj;
  met[ps + o] = d;
  met[2 * ps + o] = -b;
  met[3 * ps + o] = -c;
  met[4 * ps + o] = a;
}

hipError_t launch_metrics_finish(const double* dx_dxi, const double* dx_deta, const double* dy_dxi, const double* dy_deta,
                                 double* metrics, int64_t ps, int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  metrics_finish_kernel<<<grid, 256, 0, stream>>>(dx_dxi, dx_deta, dy_dxi, dy_deta, metrics, ps, ld, ny, nx, err);
  return hipGetLastError();
}

// Metric pairs for the stage kernel (BlockDev::mpair) from the 5 metric
// planes: (J·ξx, J·ξy) and (J·ηx, J·ηy) of a point side by side.
__global__ __launch_bounds__(256) void metric_pairs_kernel(const double* __restrict__ met, int64_t ps, int64_t ld, int ny,
                                                           int nx, double* __restrict__ pairs) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  const int64_t o = (int64_t)y * ld + x;
  pairs[2 * o] = met[ps + o];
  pairs[2 * o + 1] = met[2 * ps + o];
  pairs[2 * ps + 2 * o] = met[3 * ps + o];
  pairs[2 * ps + 2 * o + 1] = met[4 * ps + o];
}

hipError_t launch_metric_pairs(const double* metrics, int64_t ps, int64_t ld, int ny, int nx, double* pairs,
                               hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  metric_pairs_kernel<<<grid, 256, 0, stream>>>(metrics, ps, ld, ny, nx, pairs);
  return hipGetLastError();
}

// Characteristic SATs of the listed block edges at one stage (SAT_north /
// south / east / west on boundary_extracts' data, euler/src/lib.rs:1002-1144
// and :869-915), one thread per edge point: B.sat[dir] = hi·τ·T((L + σ|L|)∘
// (TI(q - g))) with q the stage input at the edge point and g its boundary
// data (a neighbour's edge, interpolated or pieced together, or far-field
// data evaluated here at the stage time). The stage kernel's edge tiles read
// them when the group's schedule evaluates the SATs here rather than in the
// tiles (a long far-field chain of pow and exp then runs on the whole GPU
// before the tiles, instead of inside them). Grid: (segment, points / 256).
__global__ __launch_bounds__(256) void edge_sat_kernel(const StageArgs A, const SatSeg* __restrict__ segs, const int mode) {
  const SatSeg sg = segs[blockIdx.x];
  const BlockDev& B = A.blocks[sg.block];
  const int dir = sg.dir;
  const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
  const int len = ns ? B.nx : B.ny;
  const int t = blockIdx.y * 256 + threadIdx.x;
  if (t >= len) return;
  const int x = ns ? t : (dir == SBP_EAST ? B.nx - 1 : 0);
  const int y = ns ? (dir == SBP_NORTH ? B.ny - 1 : 0) : t;
  const uint32_t o = (uint32_t)(y * B.ld + x);
  const uint32_t ps = (uint32_t)B.ps;
  unsigned err = 0;
  double q[4], g[4];
  load_stage_state(A, B, mode, o, q);
  edge_data(A, B, dir, t, o, mode, g, err);
  // (σ, τ): north (-1, +1), south (+1, -1), east (-1, +1), west (+1, -1);
  // north/south use J·∇η, east/west J·∇ξ
  const double sign = (dir == SBP_NORTH || dir == SBP_EAST) ? -1.0 : 1.0;
  const int m0 = ns ? 3 : 1;
  double sat[4] = {0.0, 0.0, 0.0, 0.0};
  sat_point(sat, q, g, ns ? B.hi_eta : B.hi_xi, sign, -sign, ldg(B.metrics, o), ldg(B.metrics + m0 * ps, o),
            ldg(B.metrics + (m0 + 1) * ps, o), A.gamma);
  double* out = B.sat[dir];
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c * len + t] = sat[c];
  if (err) record_error(A.error, err);
}

// Boundary data of the listed Eval edges at the stage time of `a` into
// BlockDev::edata, before the stage kernel's edge tiles read it: the expression
// interpreter's stack lives in this kernel's registers and scratch, not in
// the stage kernel's. Grid: (segment, points / 256).
__global__ __launch_bounds__(256) void eval_edges_kernel(const StageArgs A, const SatSeg* __restrict__ segs) {
  const SatSeg sg = segs[blockIdx.x];
  const BlockDev& B = A.blocks[sg.block];
  const int dir = sg.dir;
  const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
  const int len = ns ? B.nx : B.ny;
  const int t = blockIdx.y * 256 + threadIdx.x;
  if (t >= len) return;
  const int x = ns ? t : (dir == SBP_EAST ? B.nx - 1 : 0);
  const int y = ns ? (dir == SBP_NORTH ? B.ny - 1 : 0) : t;
  unsigned err = 0;
  double g[4];
  eval_edge_point(A, B, dir, (uint32_t)(y * B.ld + x), g, err);
  double* out = B.edata[dir];
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c * len + t] = g[c];
  if (err) record_error(A.error, err);
}

// Expressions on a whole block at time t into 4 planes (initial conditions
// from expressions, BaseSystem::create, multigrid/src/system.rs:143-152; the
// reference field of System::error :618-635).
__global__ __launch_bounds__(256) void eval_field_kernel(const expr::Program* __restrict__ P, double t,
                                                         const double* __restrict__ xc, const double* __restrict__ yc,
                                                         double* __restrict__ out, int64_t ps, int64_t ld, int ny, int nx,
                                                         unsigned* err) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  const int64_t o = (int64_t)y * ld + x;
  double q[4];
  if (!expr::eval_point(P, t, xc[o], yc[o], q)) {
    record_error(err, ERRBIT_EXPR);
    q[0] = q[1] = q[2] = q[3] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c * ps + o] = q[c];
}

hipError_t launch_eval_field(const expr::Program* prog, double t, const double* xc, const double* yc, double* out, int64_t ps,
                             int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  eval_field_kernel<<<grid, 256, 0, stream>>>(prog, t, xc, yc, out, ps, ld, ny, nx, err);
  return hipGetLastError();
}

hipError_t launch_eval_edges(const StageArgs& a, const SatSeg* segs, int nsegs, int maxlen, hipStream_t stream) {
  if (nsegs <= 0) return hipSuccess;
  dim3 grid(nsegs, (maxlen + 255) / 256);
  eval_edges_kernel<<<grid, 256, 0, stream>>>(a, segs);
  return hipGetLastError();
}

hipError_t launch_edge_sat(const StageArgs& a, const SatSeg* segs, int nsegs, int maxlen, int mode, hipStream_t stream) {
  if (nsegs <= 0) return hipSuccess;
  dim3 grid(nsegs, (maxlen + 255) / 256);
  edge_sat_kernel<<<grid, 256, 0, stream>>>(a, segs, mode);
  return hipGetLastError();
}

// Mixed form: the reference face terms of each listed tile of a TX x TY
// tiling (mix_ref_terms) from the state stage `mode` reads at the tile's
// centre point, one thread per tile, into StageArgs::mixref. The stage
// kernels of the step (or of the right-hand-side launch) read them instead
// of evaluating them in every thread of the tile.
__global__ __launch_bounds__(256) void mix_ref_kernel(const StageArgs A, const int mode, const int TX, const int TY,
                                                      const int ntiles) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ntiles) return;
  const TileDev T = A.tiles[i];
  const BlockDev& B = A.blocks[T.block];
  // tile origin as the stage kernel places it (tile_place; the closure
  // width only moves the written range, not the origin)
  int x0, y0, w0, w1;
  tile_place(T.tx, B.ntx, B.nx, TX, 0, x0, w0, w1);
  tile_place(T.ty, B.nty, B.ny, TY, 0, y0, w0, w1);
  const uint32_t oc = (uint32_t)(min(y0 + TY / 2, B.ny - 1) * B.ld + min(x0 + TX / 2, B.nx - 1));
  double q[4], r[16];
  load_stage_state(A, B, mode, oc, q);
  const d2* mp = reinterpret_cast<const d2*>(B.mpair);
  const d2 mx = mp[oc], me = mp[B.ps + oc];
  mix_ref_terms(q, mx.x, mx.y, me.x, me.y, A.gamma, r);
  double* out = A.mixref + 16 * (B.tile_off + (int64_t)T.ty * B.ntx + T.tx);
#pragma unroll
  for (int c = 0; c < 16; ++c) out[c] = r[c];
}

hipError_t launch_mix_ref(const StageArgs& a, int shape, int mode, int ntiles, hipStream_t stream) {
  if (ntiles <= 0) return hipSuccess;
  if (shape < 0 || shape >= kNumTileShapes || !a.mixref) return hipErrorInvalidValue;
  const TileShape& ts = kTileShapes[shape];
  mix_ref_kernel<<<(ntiles + 255) / 256, 256, 0, stream>>>(a, mode, ts.tx, ts.ty, ntiles);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Vortex initial state (Field::vortex, lib.rs:319-328).
__global__ __launch_bounds__(256) void vortex_kernel(const VortexDev vp, double gamma, double t, const double* __restrict__ xc,
                                                     const double* __restrict__ yc, double* __restrict__ out, int64_t ps,
                                                     int64_t ld, int ny, int nx, unsigned* err) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  const int64_t o = (int64_t)y * ld + x;
  double q[4] = {out[o], out[ps + o], out[2 * ps + o], out[3 * ps + o]};
  unsigned e = 0;
  vortex_point(vp, gamma, t, xc[o], yc[o], q, e);
  if (e) record_error(err, e);
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c * ps + o] = q[c];
}

hipError_t launch_vortex(const VortexDev& vp, double gamma, double t, const double* xc, const double* yc, double* out,
                         int64_t ps, int64_t ld, int ny, int nx, unsigned* err, hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  vortex_kernel<<<grid, 256, 0, stream>>>(vp, gamma, t, xc, yc, out, ps, ld, ny, nx, err);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Compensated reductions. Each thread keeps a Neumaier sum over the points it
// visits (a grid of at most kReduceMaxWG workgroups strides over the rows);
// a wavefront combines its 64 lanes with cross-lane shuffles and error-free
// TwoSum merges, the four wavefronts of a workgroup meet in LDS once, and a
// one-workgroup pass merges the per-workgroup partials into the result.
// Thread 0 of a 256-thread workgroup ends with the merge of all its threads.
__device__ __forceinline__ KSum block_ksum(KSum a) {
  __shared__ KSum wsum[4];
  a = wave_ksum(a);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    a = wsum[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) ksum_merge(a, wsum[w]);
  }
  return a;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

int reduce_workgroups(int ny) { return ny < kReduceMaxWG ? ny : kReduceMaxWG; }

template <class OX, class OY>
__global__ __launch_bounds__(256) void h2_partial_kernel(const H2Args R, double* __restrict__ partials) {
  KSum acc = {0.0, 0.0};
  for (int y = blockIdx.x; y < R.ny; y += gridDim.x) {
    for (int x = threadIdx.x; x < R.nx; x += 256) {
      const int64_t oa = (int64_t)y * R.lda + x;
      double ref[4] = {0.0, 0.0, 0.0, 0.0};
      if (R.use_vortex) {
        const int64_t oc = (int64_t)y * R.ldc + x;
        unsigned e = 0;
        vortex_point(R.vp, R.gamma, R.t, R.xc[oc], R.yc[oc], ref, e);
      } else if (R.b) {
        const int64_t ob = (int64_t)y * R.ldb + x;
#pragma unroll
        for (int c = 0; c < 4; ++c) ref[c] = R.b[c * R.psb + ob];
      }
      const double w = hweight<OX, OY>(x, R.nx, y, R.ny);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double d = R.a[c * R.psa + oa] - ref[c];
        ksum_add(acc, d * d * w);
      }
    }
  }
  acc = block_ksum(acc);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = acc.s;
    partials[2 * blockIdx.x + 1] = acc.c;
  }
}

// One workgroup per block: the block's per-tile (sum, compensation) pairs of
// advance_adaptive's error, merged into out[2b], out[2b + 1].
__global__ __launch_bounds__(256) void block_sums_kernel(const BlockDev* __restrict__ blocks,
                                                         const double* __restrict__ tile_sums, double* __restrict__ out) {
  const BlockDev& B = blocks[blockIdx.x];
  const int64_t n = (int64_t)B.ntx * B.nty;
  const double* p = tile_sums + 2 * B.tile_off;
  KSum acc = {0.0, 0.0};
  for (int64_t i = threadIdx.x; i < n; i += 256) ksum_merge(acc, KSum{p[2 * i], p[2 * i + 1]});
  acc = block_ksum(acc);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = acc.s;
    out[2 * blockIdx.x + 1] = acc.c;
  }
}

hipError_t launch_block_sums(const BlockDev* blocks, int nblocks, const double* tile_sums, double* out, hipStream_t stream) {
  if (nblocks <= 0) return hipSuccess;
  block_sums_kernel<<<nblocks, 256, 0, stream>>>(blocks, tile_sums, out);
  return hipGetLastError();
}

// One workgroup: the n per-workgroup (sum, compensation) pairs into
// scratch[2·kReduceMaxWG], scratch[2·kReduceMaxWG + 1].
__global__ __launch_bounds__(256) void h2_finish_kernel(double* __restrict__ scratch, int n) {
  KSum acc = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += 256) ksum_merge(acc, KSum{scratch[2 * i], scratch[2 * i + 1]});
  acc = block_ksum(acc);
  if (threadIdx.x == 0) {
    scratch[2 * kReduceMaxWG] = acc.s;
    scratch[2 * kReduceMaxWG + 1] = acc.c;
  }
}

template <class OX>
static hipError_t h2_partial_x(int op_eta, const H2Args& R, double* scratch, hipStream_t stream) {
  const int nwg = reduce_workgroups(R.ny);
#define SBP_H2_CASE(ID, T)                                                      \
  case ID:                                                                       \
    h2_partial_kernel<OX, ops::T><<<nwg, 256, 0, stream>>>(R, scratch);          \
    break;
  switch (op_eta) {
    SBP_H2_CASE(UPWIND4, Upwind4)
    SBP_H2_CASE(UPWIND4H2, Upwind4h2)
    SBP_H2_CASE(UPWIND9, Upwind9)
    SBP_H2_CASE(UPWIND9H2, Upwind9h2)
    SBP_H2_CASE(SBP4, SBP4)
    SBP_H2_CASE(SBP8, SBP8)
    default: return hipErrorInvalidValue;
  }
#undef SBP_H2_CASE
  return hipGetLastError();
}

hipError_t launch_h2_sum(int op_eta, int op_xi, const H2Args& R, double* scratch, hipStream_t stream) {
  if (R.ny <= 0 || R.nx <= 0) return hipErrorInvalidValue;
  hipError_t e;
  switch (op_xi) {
    case UPWIND4: e = h2_partial_x<ops::Upwind4>(op_eta, R, scratch, stream); break;
    case UPWIND4H2: e = h2_partial_x<ops::Upwind4h2>(op_eta, R, scratch, stream); break;
    case UPWIND9: e = h2_partial_x<ops::Upwind9>(op_eta, R, scratch, stream); break;
    case UPWIND9H2: e = h2_partial_x<ops::Upwind9h2>(op_eta, R, scratch, stream); break;
    case SBP4: e = h2_partial_x<ops::SBP4>(op_eta, R, scratch, stream); break;
    case SBP8: e = h2_partial_x<ops::SBP8>(op_eta, R, scratch, stream); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  h2_finish_kernel<<<1, 256, 0, stream>>>(scratch, reduce_workgroups(R.ny));
  return hipGetLastError();
}

// max |û|, |v̂| (system.rs:561-578): per-thread maxima over the strided
// points, a wavefront maximum by cross-lane shuffles, one LDS step across
// the four wavefronts, then one workgroup over the partials.
__global__ __launch_bounds__(256) void maxspeed_kernel(const double* __restrict__ q, const double* __restrict__ met, int64_t ps,
                                                       int64_t ld, int ny, int nx, double* __restrict__ partials) {
  __shared__ double wmax[2][4];
  double mu = 0.0, mv = 0.0;
  for (int y = blockIdx.x; y < ny; y += gridDim.x) {
    for (int x = threadIdx.x; x < nx; x += 256) {
      const int64_t o = (int64_t)y * ld + x;
      const double u = q[ps + o] / q[o], v = q[2 * ps + o] / q[o];
      const double uh = met[ps + o] * u + met[2 * ps + o] * v;
      const double vh = met[3 * ps + o] * u + met[4 * ps + o] * v;
      mu = fmax(mu, fabs(uh));
      mv = fmax(mv, fabs(vh));
    }
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  if ((threadIdx.x & 63) == 0) {
    wmax[0][threadIdx.x >> 6] = mu;
    wmax[1][threadIdx.x >> 6] = mv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = fmax(fmax(wmax[0][0], wmax[0][1]), fmax(wmax[0][2], wmax[0][3]));
    partials[2 * blockIdx.x + 1] = fmax(fmax(wmax[1][0], wmax[1][1]), fmax(wmax[1][2], wmax[1][3]));
  }
}

__global__ __launch_bounds__(256) void max_finish_kernel(double* __restrict__ scratch, int n) {
  __shared__ double wmax[2][4];
  double mu = 0.0, mv = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    mu = fmax(mu, scratch[2 * i]);
    mv = fmax(mv, scratch[2 * i + 1]);
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  if ((threadIdx.x & 63) == 0) {
    wmax[0][threadIdx.x >> 6] = mu;
    wmax[1][threadIdx.x >> 6] = mv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    scratch[2 * kReduceMaxWG] = fmax(fmax(wmax[0][0], wmax[0][1]), fmax(wmax[0][2], wmax[0][3]));
    scratch[2 * kReduceMaxWG + 1] = fmax(fmax(wmax[1][0], wmax[1][1]), fmax(wmax[1][2], wmax[1][3]));
  }
}

hipError_t launch_maxspeed(const double* q, const double* metrics, int64_t ps, int64_t ld, int ny, int nx, double* scratch,
                           hipStream_t stream) {
  if (ny <= 0 || nx <= 0) return hipErrorInvalidValue;
  const int nwg = reduce_workgroups(ny);
  maxspeed_kernel<<<nwg, 256, 0, stream>>>(q, metrics, ps, ld, ny, nx, scratch);
  max_finish_kernel<<<1, 256, 0, stream>>>(scratch, nwg);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Characteristic SAT of one block edge added to k (SAT_north / SAT_south /
// SAT_east / SAT_west, euler/src/lib.rs:1025-1144, SAT_characteristic
// :1150-1252), one thread per edge point: k += hi·τ·T((L + σ|L|)∘(TI(q - g))).
// k, y: 4 planes (plane stride ps, row pitch ld); metrics: 5 planes;
// g: 4 x n component-major boundary data.
__global__ __launch_bounds__(256) void sat_edge_kernel(const SatEdgeArgs A) {
  const bool ns = A.dir == SBP_NORTH || A.dir == SBP_SOUTH;
  const int n = ns ? A.nx : A.ny;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int x = ns ? t : (A.dir == SBP_EAST ? A.nx - 1 : 0);
  const int y = ns ? (A.dir == SBP_NORTH ? A.ny - 1 : 0) : t;
  const int64_t o = (int64_t)y * A.ld + x;
  double q[4], g[4], k[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    q[c] = A.y[c * A.ps + o];
    g[c] = A.g[c * n + t];
    k[c] = A.k[c * A.ps + o];
  }
  // (σ, τ): north (-1, +1), south (+1, -1), east (-1, +1), west (+1, -1)
  const double sign = (A.dir == SBP_NORTH || A.dir == SBP_EAST) ? -1.0 : 1.0;
  const double tau = -sign;
  const int m0 = ns ? 3 : 1;  // J·ηx, J·ηy for north/south, J·ξx, J·ξy for east/west
  sat_point(k, q, g, A.hi, sign, tau, A.met[o], A.met[m0 * A.ps + o], A.met[(m0 + 1) * A.ps + o], A.gamma);
#pragma unroll
  for (int c = 0; c < 4; ++c) A.k[c * A.ps + o] = k[c];
}

hipError_t launch_sat_edge(const SatEdgeArgs& a, hipStream_t stream) {
  const int n = (a.dir == SBP_NORTH || a.dir == SBP_SOUTH) ? a.nx : a.ny;
  if (n <= 0) return hipErrorInvalidValue;
  sat_edge_kernel<<<(n + 255) / 256, 256, 0, stream>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ void advance_time_kernel(double* time, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *time += dt;
}

hipError_t launch_advance_time(double* time, double dt, hipStream_t stream) {
  advance_time_kernel<<<1, 64, 0, stream>>>(time, dt);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void copy_planes_kernel(const double* __restrict__ src, int64_t sps, int64_t sld,
                                                          double* __restrict__ dst, int64_t dps, int64_t dld, int nplanes,
                                                          int ny, int nx) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  for (int p = 0; p < nplanes; ++p) dst[p * dps + (int64_t)y * dld + x] = src[p * sps + (int64_t)y * sld + x];
}

hipError_t launch_copy_planes(const double* src, int64_t sps, int64_t sld, double* dst, int64_t dps, int64_t dld, int nplanes,
                              int ny, int nx, hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  copy_planes_kernel<<<grid, 256, 0, stream>>>(src, sps, sld, dst, dps, dld, nplanes, ny, nx);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// InterpolationOperator::fine2coarse / coarse2fine (interpolation.rs:15-50)
// of one vector, one thread per output point. The table is looked up on the
// device: its pointers must be the device copies of the coefficient arrays.
__global__ __launch_bounds__(256) void interpolate_kernel(int interp_op, int fine2coarse, const double* __restrict__ in,
                                                          int nin, double* __restrict__ out, int nout) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nout) return;
  const interp::Table T = interp::table(interp_op, fine2coarse != 0);
  double v[1];
  interp::value<1>(T, i, nout, nin, [&](int m, double w[1]) { w[0] = in[m]; }, v);
  out[i] = v[0];
}

hipError_t launch_interpolate(int interp_op, bool fine2coarse, const double* in, int nin, double* out, int nout,
                              hipStream_t stream) {
  interpolate_kernel<<<(nout + 255) / 256, 256, 0, stream>>>(interp_op, fine2coarse ? 1 : 0, in, nin, out, nout);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Linear combination of an embedded RK step (clone_from + scaled_add in
// order, integrate/src/lib.rs:187-211; Field::scaled_add euler/src/lib.rs:205-212):
// each term is added with two roundings, as ndarray does.
__global__ __launch_bounds__(256) void lincomb_kernel(double* out, const double* prev, LinComb lc, int64_t ps,
                                                      int64_t ld, int nx, int64_t npts) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= 4 * npts) return;
  const int64_t c = g / npts, r = g - c * npts;
  const int64_t o = c * ps + (r / nx) * ld + (r % nx);
  double v = prev[o];
  for (int j = 0; j < lc.nk; ++j) v = v + lc.s[j] * lc.k[j][o];
  out[o] = v;
}

hipError_t launch_lincomb(double* out, const double* prev, const LinComb& lc, int64_t ps, int64_t ld, int ny, int nx,
                          hipStream_t stream) {
  const int64_t npts = (int64_t)nx * ny;
  lincomb_kernel<<<(unsigned)((4 * npts + 255) / 256), 256, 0, stream>>>(out, prev, lc, ps, ld, nx, npts);
  return hipGetLastError();
}

}  // namespace sbp
