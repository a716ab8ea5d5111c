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
// UpwindOperator2d::dissxi/disseta; algos.rs:101-187 and :371-519). One
// thread per output point; along ξ the taps are neighbours in the row, along
// η they are the rows above and below (coalesced across the wavefront).
template <class OP, int KIND, int AXIS>
__global__ __launch_bounds__(256) void apply_op_kernel(const double* __restrict__ prev, int64_t ldp, double* __restrict__ fut,
                                                       int64_t ldf, int ny, int nx, double idx) {
  using B = Blk<OP, KIND>;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= nx || y >= ny) return;
  double s;
  if (AXIS == 0)
    s = row_sum<B>(prev + (int64_t)y * ldp, 1, x, nx, 0);
  else
    s = row_sum<B>(prev + x, (int)ldp, y, ny, 0);
  fut[(int64_t)y * ldf + x] = s * idx;
}

template <class OP>
static hipError_t apply_op_t(int kind, int axis, const double* prev, int64_t ldp, double* fut, int64_t ldf, int ny, int nx,
                             hipStream_t stream) {
  const int n = axis == 0 ? nx : ny;
  const double dx = OP::is_h2 ? 1.0 / (double)(n - 2) : 1.0 / (double)(n - 1);
  const double idx = 1.0 / dx;
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  if (kind == SBP_DIFF) {
    if (axis == 0)
      apply_op_kernel<OP, SBP_DIFF, 0><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
    else
      apply_op_kernel<OP, SBP_DIFF, 1><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
  } else {
    if constexpr (OP::is_upwind) {
      if (axis == 0)
        apply_op_kernel<OP, SBP_DISS, 0><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
      else
        apply_op_kernel<OP, SBP_DISS, 1><<<grid, 256, 0, stream>>>(prev, ldp, fut, ldf, ny, nx, idx);
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

// Far-field (Vortex) boundary data of every vortex edge of a system at the
// stage's time (t + c_s·dt, times the block's step scale), evaluated once per
// stage before the edge tiles read it (BlockDev::vedge; boundary_extracts'
// Vortex case, euler/src/lib.rs:983-998): one y-block of the grid per edge.
__global__ __launch_bounds__(256) void vortex_edges_kernel(const StageArgs A, const VortexSeg* __restrict__ segs) {
  const VortexSeg sg = segs[blockIdx.y];
  const BlockDev& B = A.blocks[sg.block];
  const int dir = sg.dir;
  const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
  const int len = ns ? B.nx : B.ny;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= len) return;
  const int x = ns ? t : (dir == SBP_EAST ? B.nx - 1 : 0);
  const int y = ns ? (dir == SBP_NORTH ? B.ny - 1 : 0) : t;
  const int64_t o = (int64_t)y * B.ld + x;
  const double time = ((A.time ? *A.time : 0.0) + A.time_offset) * B.dt_scale;
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  unsigned e = 0;
  vortex_point(B.vortex[dir], A.gamma, time, B.xc[o], B.yc[o], g, e);
  if (e) record_error(A.error, e);
  double* out = B.vedge[dir];
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c * len + t] = g[c];
}

hipError_t launch_vortex_edges(const StageArgs& a, const VortexSeg* segs, int nsegs, int maxlen, hipStream_t stream) {
  dim3 grid((maxlen + 255) / 256, nsegs);
  vortex_edges_kernel<<<grid, 256, 0, stream>>>(a, segs);
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
// Compensated reductions. Each thread keeps a Neumaier sum; the workgroup
// combines lanes with error-free TwoSum steps through LDS and writes one
// (sum, compensation) pair. The host adds the pairs in long double.
struct KSum {
  double s, c;
};

__device__ __forceinline__ void ksum_add(KSum& a, double v) {
  const double t = a.s + v;
  if (fabs(a.s) >= fabs(v))
    a.c += (a.s - t) + v;
  else
    a.c += (v - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ void ksum_merge(KSum& a, const KSum& b) {
  ksum_add(a, b.s);
  a.c += b.c;
}

template <class OX, class OY>
__device__ __forceinline__ double hweight(int i, int n, int j, int m) {
  // h2_err weights (lib.rs:389-421): [H, 1.., rev(H)] scaled by 1/(n-1) (1/(n-2) for h2)
  double hx = 1.0, hy = 1.0;
  if (i < OX::M) hx = OX::h[i];
  else if (i >= n - OX::M) hx = OX::h[n - 1 - i];
  if (j < OY::M) hy = OY::h[j];
  else if (j >= m - OY::M) hy = OY::h[m - 1 - j];
  const double fx = OX::is_h2 ? 1.0 / (double)(n - 2) : 1.0 / (double)(n - 1);
  const double fy = 1.0 / (OY::is_h2 ? (double)(m - 2) : (double)(m - 1));
  return (hx * fx) * (hy * fy);
}

constexpr int kReduceRows = 8;

int reduce_h2_workgroups(int ny, int nx) { return ((nx + 255) / 256) * ((ny + kReduceRows - 1) / kReduceRows); }

template <class OX, class OY>
__global__ __launch_bounds__(256) void h2_reduce_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                        const VortexDev* vp, double gamma, double t,
                                                        const double* __restrict__ xc, const double* __restrict__ yc,
                                                        int64_t ps, int64_t ld, int ny, int nx, double* partials) {
  __shared__ KSum red[256];
  const int nbx = (nx + 255) / 256;
  const int x = (blockIdx.x % nbx) * 256 + threadIdx.x;
  const int yb = (blockIdx.x / nbx) * kReduceRows;
  KSum acc = {0.0, 0.0};
  if (x < nx) {
    for (int y = yb; y < min(yb + kReduceRows, ny); ++y) {
      const int64_t o = (int64_t)y * ld + x;
      double ref[4] = {0.0, 0.0, 0.0, 0.0};
      if (vp) {
        unsigned e = 0;
        vortex_point(*vp, gamma, t, xc[o], yc[o], ref, e);
      } else if (b) {
#pragma unroll
        for (int c = 0; c < 4; ++c) ref[c] = b[c * ps + o];
      }
      const double w = hweight<OX, OY>(x, nx, y, ny);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double d = a[c * ps + o] - ref[c];
        ksum_add(acc, d * d * w);
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) ksum_merge(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = red[0].s;
    partials[2 * blockIdx.x + 1] = red[0].c;
  }
}

template <class OX>
static hipError_t h2_reduce_x(int op_eta, const double* a, const double* b, const VortexDev* vp, double gamma, double t,
                              const double* xc, const double* yc, int64_t ps, int64_t ld, int ny, int nx, double* partials,
                              hipStream_t stream) {
  const int nwg = reduce_h2_workgroups(ny, nx);
#define SBP_H2_CASE(ID, T)                                                                                      \
  case ID:                                                                                                       \
    h2_reduce_kernel<OX, ops::T><<<nwg, 256, 0, stream>>>(a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials); \
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

hipError_t launch_h2_reduce(int op_eta, int op_xi, const double* a, const double* b, const VortexDev* vp, double gamma,
                            double t, const double* xc, const double* yc, int64_t ps, int64_t ld, int ny, int nx,
                            double* partials, hipStream_t stream) {
  switch (op_xi) {
    case UPWIND4: return h2_reduce_x<ops::Upwind4>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    case UPWIND4H2: return h2_reduce_x<ops::Upwind4h2>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    case UPWIND9: return h2_reduce_x<ops::Upwind9>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    case UPWIND9H2: return h2_reduce_x<ops::Upwind9h2>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    case SBP4: return h2_reduce_x<ops::SBP4>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    case SBP8: return h2_reduce_x<ops::SBP8>(op_eta, a, b, vp, gamma, t, xc, yc, ps, ld, ny, nx, partials, stream);
    default: return hipErrorInvalidValue;
  }
}

// max |û|, |v̂| (system.rs:561-578); wavefront reduction with DPP-free shuffles.
int reduce_max_workgroups(int ny, int nx) { return ((nx + 255) / 256) * ((ny + kReduceRows - 1) / kReduceRows); }

__global__ __launch_bounds__(256) void maxspeed_kernel(const double* __restrict__ q, const double* __restrict__ met, int64_t ps,
                                                       int64_t ld, int ny, int nx, double* partials) {
  __shared__ double red[2][256];
  const int nbx = (nx + 255) / 256;
  const int x = (blockIdx.x % nbx) * 256 + threadIdx.x;
  const int yb = (blockIdx.x / nbx) * kReduceRows;
  double mu = 0.0, mv = 0.0;
  if (x < nx) {
    for (int y = yb; y < min(yb + kReduceRows, ny); ++y) {
      const int64_t o = (int64_t)y * ld + x;
      const double u = q[ps + o] / q[o], v = q[2 * ps + o] / q[o];
      const double uh = met[ps + o] * u + met[2 * ps + o] * v;
      const double vh = met[3 * ps + o] * u + met[4 * ps + o] * v;
      mu = fmax(mu, fabs(uh));
      mv = fmax(mv, fabs(vh));
    }
  }
  red[0][threadIdx.x] = mu;
  red[1][threadIdx.x] = mv;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + off]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = red[0][0];
    partials[2 * blockIdx.x + 1] = red[1][0];
  }
}

hipError_t launch_maxspeed(const double* q, const double* metrics, int64_t ps, int64_t ld, int ny, int nx, double* partials,
                           hipStream_t stream) {
  maxspeed_kernel<<<reduce_max_workgroups(ny, nx), 256, 0, stream>>>(q, metrics, ps, ld, ny, nx, partials);
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
