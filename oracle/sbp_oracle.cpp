// SummationByParts Euler RK4 oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain-C++ CPU restatement of the reference's Euler time-step path. The
// parity tests compare the HIP library against it, and bench.py times it as
// the CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.
// The product (summationbyparts_amd) never links, imports or calls it.
//
// Compiled with -ffp-contract=off so every a*b+c is two roundings, exactly
// where the reference's Rust code is (Rust never contracts), and std::fma only
// where the reference calls `mul_add`.
//
// What follows the reference, function by function:
//   diff_rows      sbp/src/operators/algos.rs:101-187 (row slices, matmul
//                  order utils/constmatrix/src/lib.rs:118-147, no FMA)
//   diff_cols      sbp/src/operators/algos.rs:371-519 (η via transposed view:
//                  8-lane SIMD part with fused mul_add from 0, scalar remainder)
//   diff_2d        dispatch sbp/src/operators.rs:55-106, algos.rs:539-555
//   metrics        sbp/src/grid.rs:95-139
//   vortex         euler/src/vortex.rs:23-113
//   fluxes         euler/src/lib.rs:680-735 (+ pressure :478-480)
//   dissipation    euler/src/lib.rs:619-674
//   rhs            euler/src/lib.rs:483-515 (trad), :571-616 (upwind)
//   sat            euler/src/lib.rs:1002-1252
//   boundaries     euler/src/lib.rs:763-802, :804-915 (This / Grid / Vortex /
//                  Interpolate / MultiGrid / Eval, the last through a callback)
//   interpolate    sbp/src/operators/interpolation.rs:15-50 (2:1 interfaces;
//                  tables in sbp_interp_tables.h)
//   integrate      utils/integrate/src/lib.rs:30-35, :176-222 (Rk4; any
//                  explicit tableau given as arrays: system_step_tableau)
//   rhs_no_sat     euler/src/lib.rs:518-568 (block_rhs with the SATs off)
//   sat_edge       euler/src/lib.rs:1025-1144 (one SAT function on arrays)
//   embedded RK    utils/integrate/src/lib.rs:138-151, :176-251
//                  (BogackiShampine, integrate_embedded_rk)
//   adaptive       euler/src/lib.rs:149-185 (advance_adaptive)
//   h2_err         euler/src/lib.rs:377-432
//   max_dt         multigrid/src/system.rs:542-589
//   error          multigrid/src/system.rs:618-635
// Parity pins for each function are listed in tests/test_oracle.py.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "sbp_interp_tables.h"
#include "sbp_tables.h"

namespace {

enum { OP_UPWIND4 = 0, OP_UPWIND4H2 = 1, OP_UPWIND9 = 2, OP_UPWIND9H2 = 3, OP_SBP4 = 4, OP_SBP8 = 5, OP_COUNT = 6 };
enum { KIND_DIFF = 0, KIND_DISS = 1 };
enum { ST_OK = 0, ST_INVALID = 1, ST_SHAPE = 2, ST_DETJ = 3, ST_RHO = 4, ST_PRESSURE = 5, ST_GAMMA = 6, ST_EDGE = 7,
       ST_NOT_UPWIND = 9 };
enum { BC_THIS = 0, BC_GRID = 1, BC_VORTEX = 2, BC_INTERPOLATE = 3, BC_MULTIGRID = 4, BC_EVAL = 5 };

// Eval boundary data (lib.rs:853-866): the evaluator is a callback that fills
// rho, rhou, rhov, e (4 x n, component-major) at time t on n edge points.
typedef void (*orc_eval_fn)(double t, int n, const double* x, const double* y, double* out, void* user);
enum { IOP_4 = 0, IOP_8 = 1, IOP_9 = 2, IOP_9H2 = 3, IOP_COUNT = 4 };

// One direction (fine to coarse or coarse to fine) of an interpolation
// operator: start block, repeating diagonal rows and the jump that places
// them (interpolation4.rs:74-93 and the other three alike).
struct InterpDef {
  const double* block;
  const double* diag;
  int brows, bcols, drows, dcols, j0, j1;
  bool h2;
};

template <class I>
InterpDef interp_def_of(bool f2c) {
  if (f2c)
    return {&I::f2c_block[0][0], &I::f2c_diag[0][0], I::f2c_block_rows, I::f2c_block_cols, I::f2c_diag_rows,
            I::f2c_diag_cols, I::f2c_jump[0], I::f2c_jump[1], I::is_h2};
  return {&I::c2f_block[0][0], &I::c2f_diag[0][0], I::c2f_block_rows, I::c2f_block_cols, I::c2f_diag_rows,
          I::c2f_diag_cols, I::c2f_jump[0], I::c2f_jump[1], I::is_h2};
}

bool interp_def(int op, bool f2c, InterpDef& d) {
  using namespace interp_tables;
  switch (op) {
    case IOP_4: d = interp_def_of<Interpolation4>(f2c); return true;
    case IOP_8: d = interp_def_of<Interpolation8>(f2c); return true;
    case IOP_9: d = interp_def_of<Interpolation9>(f2c); return true;
    case IOP_9H2: d = interp_def_of<Interpolation9h2>(f2c); return true;
    default: return false;
  }
}

// fine2coarse / coarse2fine length asserts (interpolation4.rs:75,85;
// interpolation9h2.rs:53,64): fine = 2 coarse - 1, or 2 (coarse - 1) on h2 grids
// and the input indices the loop below reads (interpolation.rs:29-46 index
// `input` directly, so a short edge panics there): start and end rows read
// bcols points from either end, the last interior row dcols points from
// j0 + j1·(its row group)
bool interp_lengths(const InterpDef& d, bool f2c, int nin, int nout) {
  const int coarse = f2c ? nout : nin, fine = f2c ? nin : nout;
  if (!(d.h2 ? fine == 2 * (coarse - 1) : fine == 2 * coarse - 1)) return false;
  if (nin <= 0 || nout <= 0 || d.bcols > nin) return false;
  const int last = nout - d.brows - 1;
  if (last >= d.brows) {
    const int index = d.j0 + d.j1 * ((last - d.brows) / d.drows);
    if (index + d.dcols > nin) return false;
  }
  return true;
}

// interpolation.rs:15-50: output rows below the block height take the block,
// rows within the block height of the end take the block mirrored (row from
// the end, columns reversed, aligned with the input's end), the rest the
// diagonal row (n mod rows) at input index jump.0 + jump.1·(n div rows);
// each output is a sum from 0.0 over the taps in order, unfused.
void interpolate(const InterpDef& d, const double* in, size_t is, int nin, double* out, size_t os, int nout) {
  for (int i = 0; i < nout; ++i) {
    double acc = 0.0;
    if (i < d.brows) {
      for (int k = 0; k < d.bcols; ++k) acc += in[(size_t)k * is] * d.block[i * d.bcols + k];
    } else if (nout - i <= d.brows) {
      const int row = nout - i - 1, index = nin - d.bcols;
      for (int k = 0; k < d.bcols; ++k) acc += in[(size_t)(index + k) * is] * d.block[row * d.bcols + (d.bcols - k - 1)];
    } else {
      const int n = i - d.brows;
      const int index = d.j0 + d.j1 * (n / d.drows), row = n % d.drows;
      for (int k = 0; k < d.dcols; ++k) acc += in[(size_t)(index + k) * is] * d.diag[row * d.dcols + k];
    }
    out[(size_t)i * os] = acc;
  }
}
enum { DIR_NORTH = 0, DIR_SOUTH = 1, DIR_EAST = 2, DIR_WEST = 3 };
enum { RHS_AUTO = 0, RHS_TRAD = 1, RHS_UPWIND = 2 };

struct BlockMat {
  int M = 0, N = 0, D = 0;
  std::vector<double> start, diag, end;  // start/end: M*N row-major
  bool present = false;
};

struct OpDef {
  int M, N, D;
  bool h2, upwind;
  const double* h;
  BlockMat diff, diss;
};

BlockMat make_block(int M, int N, int D, const double* start, const double* diag, bool negate_end) {
  BlockMat b;
  b.M = M;
  b.N = N;
  b.D = D;
  b.present = true;
  b.start.assign(start, start + M * N);
  b.diag.assign(diag, diag + D);
  b.end.resize(M * N);
  // end = flip_lr(flip_ud(start)) (·-1 for DIFF): constmatrix/src/lib.rs:302-343
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double v = start[(M - 1 - i) * N + (N - 1 - j)];
      b.end[i * N + j] = negate_end ? -v : v;
    }
  return b;
}

const OpDef* get_op(int op) {
  static OpDef defs[OP_COUNT];
  static bool init = false;
  if (!init) {
    using namespace orc_tab;
    defs[OP_UPWIND4] = {4, 8, 7, false, true, up4_h, make_block(4, 8, 7, &up4_diff_start[0][0], up4_diff_diag, true),
                        make_block(4, 8, 7, &up4_diss_start[0][0], up4_diss_diag, false)};
    defs[OP_UPWIND4H2] = {4, 8, 7, true, true, up4h2_h, make_block(4, 8, 7, &up4h2_diff_start[0][0], up4h2_diff_diag, true),
                          make_block(4, 8, 7, &up4h2_diss_start[0][0], up4h2_diss_diag, false)};
    defs[OP_UPWIND9] = {8, 13, 11, false, true, up9_h, make_block(8, 13, 11, &up9_diff_start[0][0], up9_diff_diag, true),
                        make_block(8, 13, 11, &up9_diss_start[0][0], up9_diss_diag, false)};
    defs[OP_UPWIND9H2] = {8, 13, 11, true, true, up9h2_h,
                          make_block(8, 13, 11, &up9h2_diff_start[0][0], up9h2_diff_diag, true),
                          make_block(8, 13, 11, &up9h2_diss_start[0][0], up9h2_diss_diag, false)};
    defs[OP_SBP4] = {4, 8, 5, false, false, sbp4_h, make_block(4, 8, 5, &sbp4_diff_start[0][0], sbp4_diff_diag, true), BlockMat()};
    defs[OP_SBP8] = {8, 12, 9, false, false, sbp8_h, make_block(8, 12, 9, &sbp8_diff_start[0][0], sbp8_diff_diag, true), BlockMat()};
    init = true;
  }
  if (op < 0 || op >= OP_COUNT) return nullptr;
  return &defs[op];
}

const BlockMat* get_block(int op, int kind) {
  const OpDef* d = get_op(op);
  if (!d) return nullptr;
  if (kind == KIND_DIFF) return &d->diff;
  if (kind == KIND_DISS && d->upwind) return &d->diss;
  return nullptr;
}

// Threads inside one block (orc_set_inner_threads; 1 = the reference's
// one thread per block). Work is split into contiguous ranges of rows or
// points, each computed with the same arithmetic as the sequential loop, so
// the results do not depend on the thread count.
int g_inner_threads = 1;

template <class F>
void pfor(size_t n, F&& f) {
  const int T = g_inner_threads;
  if (T <= 1 || n < 256) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + (size_t)T - 1) / (size_t)T;
  for (int t = 0; t < T; ++t) {
    const size_t lo = (size_t)t * chunk, hi = std::min(n, lo + chunk);
    if (lo < hi) th.emplace_back(f, lo, hi);
  }
  for (auto& x : th) x.join();
}

inline double inv_spacing(int n, bool h2) {
  // algos.rs:125-130: dx = 1/(n-1) (1/(n-2) for h2), idx = 1/dx
  double dx = h2 ? 1.0 / (double)(n - 2) : 1.0 / (double)(n - 1);
  return 1.0 / dx;
}

// ξ-direction operator on every row of a C-order (ny, nx) array.
// algos.rs:101-187: each row is a slice; start/end via matmul_float_into
// (t = a0*b0; t = t + ak*bk), interior window dot product, then *idx.
void diff_rows(const BlockMat& B, bool h2, const double* prev, double* fut, int ny, int nx) {
  const int M = B.M, N = B.N, D = B.D;
  const double idx = inv_spacing(nx, h2);
  const int skip = M - (D - 1) / 2;
  pfor((size_t)ny, [&](size_t j0, size_t j1) {
  for (int j = (int)j0; j < (int)j1; ++j) {
    const double* p = prev + (size_t)j * nx;
    double* f = fut + (size_t)j * nx;
    for (int i = 0; i < M; ++i) {
      const double* a = &B.start[i * N];
      double t = a[0] * p[0];
      for (int k = 1; k < N; ++k) t = t + a[k] * p[k];
      f[i] = t * idx;
    }
    for (int i = M; i < nx - M; ++i) {
      const double* w = p + skip + (i - M);
      double t = B.diag[0] * w[0];
      for (int k = 1; k < D; ++k) t = t + B.diag[k] * w[k];
      f[i] = t * idx;
    }
    const double* pe = p + (nx - N);
    for (int i = 0; i < M; ++i) {
      const double* a = &B.end[i * N];
      double t = a[0] * pe[0];
      for (int k = 1; k < N; ++k) t = t + a[k] * pe[k];
      f[nx - M + i] = t * idx;
    }
  }
  });
}

// η-direction operator on a C-order (ny, nx) array. The reference applies
// diffxi to the transposed view, which lands in diff_op_2d_sliceable_y_simd
// (algos.rs:371-519): for ξ-indices in the 8-lane SIMD part the sum starts at
// 0 and each tap is a fused mul_add; the remaining ξ-indices use an unfused
// scalar sum. Output rows are η-indices.
void diff_cols(const BlockMat& B, bool h2, const double* prev, double* fut, int ny, int nx) {
  const int M = B.M, N = B.N, D = B.D;
  const double idx = inv_spacing(ny, h2);
  const int simd = 8 * (nx / 8);
  const int half = (D - 1) / 2;
  auto row = [&](int r) { return prev + (size_t)r * nx; };
  auto block_rows = [&](const std::vector<double>& mat, int in0, int out0) {
    for (int r = 0; r < M; ++r) {
      const double* a = &mat[r * N];
      double* f = fut + (size_t)(out0 + r) * nx;
      for (int i = 0; i < simd; ++i) {
        double s = 0.0;
        for (int k = 0; k < N; ++k) s = std::fma(row(in0 + k)[i], a[k], s);
        f[i] = s * idx;
      }
      for (int i = simd; i < nx; ++i) {
        double s = 0.0;
        for (int k = 0; k < N; ++k) s = s + a[k] * row(in0 + k)[i];
        f[i] = s * idx;
      }
    }
  };
  block_rows(B.start, 0, 0);
  pfor((size_t)(ny - 2 * M), [&](size_t r0, size_t r1) {
  for (int r = M + (int)r0; r < M + (int)r1; ++r) {
    double* f = fut + (size_t)r * nx;
    for (int i = 0; i < simd; ++i) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s = std::fma(row(r - half + k)[i], B.diag[k], s);
      f[i] = s * idx;
    }
    for (int i = simd; i < nx; ++i) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s = s + B.diag[k] * row(r - half + k)[i];
      f[i] = idx * s;
    }
  }
  });
  block_rows(B.end, ny - N, ny - M);
}

// axis 0 = ξ (along nx), axis 1 = η (along ny). Shape checks as algos.rs:118-121.
int diff_2d(int op, int kind, int axis, const double* prev, double* fut, int ny, int nx) {
  const OpDef* d = get_op(op);
  if (!d) return ST_INVALID;
  const BlockMat* B = get_block(op, kind);
  if (!B) return ST_NOT_UPWIND;
  int n = axis == 0 ? nx : ny;
  if (n < 2 * B->M || n < B->N) return ST_SHAPE;
  if (axis == 0)
    diff_rows(*B, d->h2, prev, fut, ny, nx);
  else
    diff_cols(*B, d->h2, prev, fut, ny, nx);
  return ST_OK;
}

double g_gamma = 1.4;
bool g_gamma_set = false;

struct Vortice {
  double x0, y0, rstar, eps;
};
}  // namespace

extern "C" {
struct orc_vortex_params {
  int32_t nvortices;
  int32_t pad_;
  Vortice v[5];
  double mach;
};
}

namespace {

// vortex.rs:26-113
int vortex_eval(const orc_vortex_params& vp, double time, const double* x, const double* y, size_t n, size_t stride_x,
                size_t stride_y, double* rho, double* rhou, double* rhov, double* e, size_t stride_out) {
  const double gamma = g_gamma;
  const double m = vp.mach;
  const double p_inf = 1.0 / (gamma * m * m);
  const double rho_inf = 1.0, u_inf = 1.0, v_inf = 0.0;
  const double e_inf = p_inf / (gamma - 1.0) + rho_inf * (u_inf * u_inf + v_inf * v_inf) / 2.0;
  const double PI = M_PI;
  for (size_t q = 0; q < n; ++q) {
    double xx = x[q * stride_x], yy = y[q * stride_y];
    double& r = rho[q * stride_out];
    double& ru = rhou[q * stride_out];
    double& rv = rhov[q * stride_out];
    double& ee = e[q * stride_out];
    if (vp.nvortices <= 0) {
      // vortex.rs:56-60 writes rhou twice and leaves rhov untouched
      r = rho_inf;
      ru = rho_inf * u_inf;
      ru = rho_inf * v_inf;
      ee = e_inf;
      continue;
    }
    for (int iv = 0; iv < vp.nvortices; ++iv) {
      const Vortice& vo = vp.v[iv];
      double rstar = vo.rstar, eps = vo.eps;
      double dx = (xx - vo.x0) - time;
      double dy = yy - vo.y0;
      double f = (1.0 - (dx * dx + dy * dy)) / (rstar * rstar);
      double rv_ = std::pow(1.0 - eps * eps * (gamma - 1.0) * m * m / (8.0 * PI * PI * p_inf * rstar * rstar) * std::exp(f),
                            1.0 / (gamma - 1.0));
      if (!(rv_ > 0.0)) return ST_RHO;
      double p = std::pow(rv_, gamma) * p_inf;
      double u = 1.0 - eps * dy / (2.0 * PI * std::sqrt(p_inf) * rstar * rstar) * std::exp(f / 2.0);
      double v = eps * dx / (2.0 * PI * std::sqrt(p_inf) * rstar * rstar) * std::exp(f / 2.0);
      if (!(p > 0.0)) return ST_PRESSURE;
      if (iv == 0) {
        r = rv_;
        ru = r * u;
        rv = r * v;
        ee = p / (gamma - 1.0) + r * (u * u + v * v) / 2.0;
      } else {
        r += rv_ - rho_inf;
        if (!(r > 0.0)) return ST_RHO;
        ru += rv_ * u - rho_inf * u_inf;
        rv += rv_ * v - rho_inf * v_inf;
        ee += (p / (gamma - 1.0) + rv_ * (u * u + v * v) / 2.0) - e_inf;
      }
    }
  }
  return ST_OK;
}

inline double pressure(double gamma, double rho, double rhou, double rhov, double e) {
  return (gamma - 1.0) * (e - (rhou * rhou + rhov * rhov) / (2.0 * rho));
}

struct Block {
  int ny, nx, op_eta, op_xi;
  std::vector<double> x, y;
  std::vector<double> detj, jxx, jxy, jyx, jyy;  // J, J·ξx, J·ξy, J·ηx, J·ηy
  int bc_kind[4] = {BC_THIS, BC_THIS, BC_THIS, BC_THIS};
  int bc_nb[4] = {0, 0, 0, 0};
  int bc_iop[4] = {0, 0, 0, 0};           // Interpolate: operator
  std::vector<int> bc_pieces[4];          // MultiGrid: (block, start, end) triples
  orc_vortex_params bc_vortex[4];
  orc_eval_fn bc_eval[4] = {nullptr, nullptr, nullptr, nullptr};  // Eval: evaluator and its user data
  void* bc_eval_user[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t npts() const { return (size_t)ny * nx; }
};

// grid.rs:95-139 with the tuple routing diffxi→op_xi, diffeta→op_eta
int compute_metrics(Block& b) {
  size_t n = b.npts();
  std::vector<double> dx_dxi(n), dx_deta(n), dy_dxi(n), dy_deta(n);
  int st;
  if ((st = diff_2d(b.op_xi, KIND_DIFF, 0, b.x.data(), dx_dxi.data(), b.ny, b.nx))) return st;
  if ((st = diff_2d(b.op_eta, KIND_DIFF, 1, b.x.data(), dx_deta.data(), b.ny, b.nx))) return st;
  if ((st = diff_2d(b.op_xi, KIND_DIFF, 0, b.y.data(), dy_dxi.data(), b.ny, b.nx))) return st;
  if ((st = diff_2d(b.op_eta, KIND_DIFF, 1, b.y.data(), dy_deta.data(), b.ny, b.nx))) return st;
  b.detj.resize(n);
  b.jxx.resize(n);
  b.jxy.resize(n);
  b.jyx.resize(n);
  b.jyy.resize(n);
  for (size_t q = 0; q < n; ++q) {
    double dj = dx_dxi[q] * dy_deta[q] - dx_deta[q] * dy_dxi[q];
    if (!(dj > 0.0)) return ST_DETJ;
    b.detj[q] = dj;
    b.jxx[q] = dy_deta[q];
    b.jxy[q] = -dx_deta[q];
    b.jyx[q] = -dy_dxi[q];
    b.jyy[q] = dx_dxi[q];
  }
  return ST_OK;
}

struct Work {
  std::vector<double> ehat, fhat, dE, dF, adx, ady, p;
  std::vector<double> edge[4];
  void ensure(size_t n, int nx, int ny) {
    if (ehat.size() != 4 * n) {
      for (auto* v : {&ehat, &fhat, &dE, &dF, &adx, &ady}) v->assign(4 * n, 0.0);
      p.assign(n, 0.0);
    }
    edge[DIR_NORTH].assign(4 * (size_t)nx, 0.0);
    edge[DIR_SOUTH].assign(4 * (size_t)nx, 0.0);
    edge[DIR_EAST].assign(4 * (size_t)ny, 0.0);
    edge[DIR_WEST].assign(4 * (size_t)ny, 0.0);
  }
};

// SAT_characteristic, lib.rs:1150-1252. k, y, z are 4 x n edge arrays with
// component stride cs and point stride ps (k, y index the block fields).
void sat_characteristic(double* k, size_t kcs, size_t kps, const double* y, size_t ycs, size_t yps, const double* z,
                        size_t zcs, size_t zps, int n, double hi, double sign, double tau, const double* detj,
                        const double* ddx, const double* ddy, size_t mps) {
  const double gamma = g_gamma;
  for (int q = 0; q < n; ++q) {
    double rho = y[0 * ycs + q * yps], rhou = y[1 * ycs + q * yps], rhov = y[2 * ycs + q * yps], e = y[3 * ycs + q * yps];
    double dj = detj[q * mps];
    double kx_ = ddx[q * mps] / dj;
    double ky_ = ddy[q * mps] / dj;
    double r = std::hypot(kx_, ky_);
    double kx = kx_ / r, ky = ky_ / r;
    double u = rhou / rho, v = rhov / rho;
    double theta = kx * u + ky * v;
    double p = pressure(gamma, rho, rhou, rhov, e);
    double c = std::sqrt(gamma * p / rho);
    double phi2 = (gamma - 1.0) * (u * u + v * v) / 2.0;
    double alpha = rho / (M_SQRT2 * c);
    double phi2_c2 = (phi2 + c * c) / (gamma - 1.0);
    double T[4][4] = {
        {1.0, 0.0, alpha, alpha},
        {u, ky, alpha * (u + kx * c), alpha * (u - kx * c)},
        {v, -kx, alpha * (v + ky * c), alpha * (v - ky * c)},
        {phi2 / (gamma - 1.0), rho * (ky * u - kx * v), alpha * (phi2_c2 + c * theta), alpha * (phi2_c2 - c * theta)},
    };
    double U = kx_ * u + ky_ * v;
    double hyp = std::hypot(kx_, ky_);
    double L[4] = {U, U, U + c * hyp, U - c * hyp};
    double beta = 1.0 / (2.0 * c * c);
    double TI[4][4] = {
        {1.0 - phi2 / (c * c), (gamma - 1.0) * u / (c * c), (gamma - 1.0) * v / (c * c), -(gamma - 1.0) / (c * c)},
        {-(ky * u - kx * v) / rho, ky / rho, -kx / rho, 0.0},
        {beta * (phi2 - c * theta), beta * (kx * c - (gamma - 1.0) * u), beta * (ky * c - (gamma - 1.0) * v), beta * (gamma - 1.0)},
        {beta * (phi2 + c * theta), -beta * (kx * c + (gamma - 1.0) * u), -beta * (ky * c + (gamma - 1.0) * v), beta * (gamma - 1.0)},
    };
    double res[4];
    for (int c4 = 0; c4 < 4; ++c4) res[c4] = y[c4 * ycs + q * yps] - z[c4 * zcs + q * zps];
    double TIres[4] = {0, 0, 0, 0};
    for (int row = 0; row < 4; ++row)
      for (int col = 0; col < 4; ++col) TIres[row] += TI[row][col] * res[col];
    double LTIres[4];
    for (int row = 0; row < 4; ++row) LTIres[row] = (L[row] + sign * std::fabs(L[row])) * TIres[row];
    double TLTIres[4] = {0, 0, 0, 0};
    for (int row = 0; row < 4; ++row)
      for (int col = 0; col < 4; ++col) TLTIres[row] += T[row][col] * LTIres[col];
    for (int c4 = 0; c4 < 4; ++c4) k[c4 * kcs + q * kps] += hi * tau * TLTIres[c4];
  }
}

struct System {
  std::vector<Block> blocks;
  std::vector<std::vector<double>> fnow, fnext;
  std::vector<std::vector<double>> k[4];
  std::vector<Work> work;
  double time = 0.0;
  int rhs_kind = RHS_AUTO;
  bool sat = true;  // false: RHS_no_SAT (lib.rs:518-568)
};

bool block_is_upwind(const Block& b) { return get_op(b.op_eta)->upwind && get_op(b.op_xi)->upwind; }

// Boundary data g for each edge (boundary_extracts, lib.rs:869-915): pointer,
// component stride and point stride into either a field or an edge buffer.
struct EdgeView {
  const double* p;
  size_t cs, ps;
};

int extract_edge(System& S, size_t bi, const std::vector<std::vector<double>>& fields, int dir, double time, EdgeView& out) {
  Block& b = S.blocks[bi];
  Work& w = S.work[bi];
  int kind = b.bc_kind[dir];
  // seldir: north takes the source's south edge, south its north, east its west, west its east
  auto sel = [&](size_t src) -> EdgeView {
    const Block& s = S.blocks[src];
    size_t n = s.npts();
    const double* f = fields[src].data();
    switch (dir) {
      case DIR_NORTH: return {f, n, 1};                                  // source south row 0
      case DIR_SOUTH: return {f + (size_t)(s.ny - 1) * s.nx, n, 1};      // source north row
      case DIR_EAST: return {f, n, (size_t)s.nx};                        // source west column
      default: return {f + (s.nx - 1), n, (size_t)s.nx};                 // source east column
    }
  };
  if (kind == BC_THIS) {
    out = sel(bi);
  } else if (kind == BC_GRID) {
    size_t src = (size_t)b.bc_nb[dir];
    if (src >= S.blocks.size()) return ST_INVALID;
    const Block& s = S.blocks[src];
    int need = (dir == DIR_NORTH || dir == DIR_SOUTH) ? b.nx : b.ny;
    int have = (dir == DIR_NORTH || dir == DIR_SOUTH) ? s.nx : s.ny;
    if (need != have) return ST_EDGE;
    out = sel(src);
  } else if (kind == BC_EVAL) {
    // Eval (lib.rs:853-866): the evaluator on this block's own edge
    // coordinates at the stage time
    const int n = (dir == DIR_NORTH || dir == DIR_SOUTH) ? b.nx : b.ny;
    std::vector<double> ex(n), ey(n);
    for (int i = 0; i < n; ++i) {
      size_t q;
      switch (dir) {
        case DIR_NORTH: q = (size_t)(b.ny - 1) * b.nx + i; break;
        case DIR_SOUTH: q = (size_t)i; break;
        case DIR_EAST: q = (size_t)i * b.nx + (b.nx - 1); break;
        default: q = (size_t)i * b.nx; break;
      }
      ex[i] = b.x[q];
      ey[i] = b.y[q];
    }
    if (!b.bc_eval[dir]) return ST_INVALID;
    double* eb = w.edge[dir].data();
    b.bc_eval[dir](time, n, ex.data(), ey.data(), eb, b.bc_eval_user[dir]);
    out = {eb, (size_t)n, 1};
  } else if (kind == BC_VORTEX) {
    // vortexify on this block's own edge coordinates (lib.rs:983-998, grid.rs:68-91)
    int n = (dir == DIR_NORTH || dir == DIR_SOUTH) ? b.nx : b.ny;
    const double *gx, *gy;
    size_t gs;
    switch (dir) {
      case DIR_NORTH: gx = &b.x[(size_t)(b.ny - 1) * b.nx]; gy = &b.y[(size_t)(b.ny - 1) * b.nx]; gs = 1; break;
      case DIR_SOUTH: gx = &b.x[0]; gy = &b.y[0]; gs = 1; break;
      case DIR_EAST: gx = &b.x[b.nx - 1]; gy = &b.y[b.nx - 1]; gs = b.nx; break;
      default: gx = &b.x[0]; gy = &b.y[0]; gs = b.nx; break;
    }
    double* eb = w.edge[dir].data();
    int st = vortex_eval(b.bc_vortex[dir], time, gx, gy, n, gs, gs, eb, eb + n, eb + 2 * n, eb + 3 * n, 1);
    if (st) return st;
    out = {eb, (size_t)n, 1};
  } else if (kind == BC_INTERPOLATE) {
    // lib.rs:816-828: the source's opposite edge interpolated onto this edge;
    // fine to coarse when this block has fewer points along ξ than the
    // source (the reference compares nx for every edge direction)
    size_t src = (size_t)b.bc_nb[dir];
    if (src >= S.blocks.size()) return ST_INVALID;
    const Block& s = S.blocks[src];
    const bool ns = dir == DIR_NORTH || dir == DIR_SOUTH;
    const int nout = ns ? b.nx : b.ny, nin = ns ? s.nx : s.ny;
    const bool f2c = b.nx < s.nx;
    InterpDef d;
    if (!interp_def(b.bc_iop[dir], f2c, d)) return ST_INVALID;
    if (!interp_lengths(d, f2c, nin, nout)) return ST_EDGE;
    const EdgeView from = sel(src);
    double* eb = w.edge[dir].data();
    for (int c = 0; c < 4; ++c) interpolate(d, from.p + c * from.cs, from.ps, nin, eb + (size_t)c * nout, 1, nout);
    out = {eb, (size_t)nout, 1};
  } else if (kind == BC_MULTIGRID) {
    // lib.rs:829-849: the pieces (block g, start, end) of the sources'
    // opposite edges are copied one after another, each starting on the last
    // point of the one before (which it overwrites); after a piece is
    // copied, its first point is halved unless it is the last piece
    const int n = (dir == DIR_NORTH || dir == DIR_SOUTH) ? b.nx : b.ny;
    const std::vector<int>& pc = b.bc_pieces[dir];
    if (pc.empty() || pc.size() % 3) return ST_EDGE;
    double* eb = w.edge[dir].data();
    int i = 0;
    size_t remaining = pc.size() / 3;
    for (size_t p = 0; p < pc.size(); p += 3) {
      const size_t g = (size_t)pc[p];
      const int start = pc[p + 1], end = pc[p + 2];
      if (g >= S.blocks.size()) return ST_INVALID;
      const Block& s = S.blocks[g];
      const int have = (dir == DIR_NORTH || dir == DIR_SOUTH) ? s.nx : s.ny;
      const int len = end - start;
      if (start < 0 || end > have || len < 1 || i + len > n) return ST_EDGE;
      const EdgeView from = sel(g);
      for (int c = 0; c < 4; ++c)
        for (int k = 0; k < len; ++k) eb[(size_t)c * n + i + k] = from.p[c * from.cs + (size_t)(start + k) * from.ps];
      remaining -= 1;
      if (remaining == 0) {
        i += len;
        if (i != n) return ST_EDGE;
      } else {
        for (int c = 0; c < 4; ++c) eb[(size_t)c * n + i] /= 2.0;
        i += len - 1;
      }
    }
    out = {eb, (size_t)n, 1};
  } else {
    return ST_INVALID;
  }
  return ST_OK;
}

// RHS for one block (RHS_trad lib.rs:483-515 / RHS_upwind :571-616) with SATs.
int block_rhs(System& S, size_t bi, const std::vector<std::vector<double>>& fields, double time, double* kout, bool upwind) {
  Block& b = S.blocks[bi];
  Work& w = S.work[bi];
  const size_t n = b.npts();
  const double* y = fields[bi].data();
  const double gamma = g_gamma;
  if (!g_gamma_set) return ST_GAMMA;
  double *eh = w.ehat.data(), *fh = w.fhat.data(), *p = w.p.data();
  const double *rho = y, *rhou = y + n, *rhov = y + 2 * n, *e = y + 3 * n;
  // fluxes, lib.rs:680-735
  pfor(n, [&](size_t q0, size_t q1) {
  for (size_t q = q0; q < q1; ++q) p[q] = pressure(gamma, rho[q], rhou[q], rhov[q], e[q]);
  for (size_t q = q0; q < q1; ++q) {
    eh[q] = rhou[q];
    eh[n + q] = rhou[q] * rhou[q] / rho[q] + p[q];
    eh[2 * n + q] = rhou[q] * rhov[q] / rho[q];
    eh[3 * n + q] = rhou[q] * (e[q] + p[q]) / rho[q];
    fh[q] = rhov[q];
    fh[n + q] = eh[2 * n + q];
    fh[2 * n + q] = rhov[q] * rhov[q] / rho[q] + p[q];
    fh[3 * n + q] = rhov[q] * (e[q] + p[q]) / rho[q];
  }
  for (int c = 0; c < 4; ++c)
    for (size_t q = q0; q < q1; ++q) {
      double ef = eh[c * n + q], ff = fh[c * n + q];
      eh[c * n + q] = b.jxx[q] * ef + b.jxy[q] * ff;
      fh[c * n + q] = b.jyx[q] * ef + b.jyy[q] * ff;
    }
  });
  int st;
  for (int c = 0; c < 4; ++c) {
    if ((st = diff_2d(b.op_xi, KIND_DIFF, 0, eh + c * n, w.dE.data() + c * n, b.ny, b.nx))) return st;
    if ((st = diff_2d(b.op_eta, KIND_DIFF, 1, fh + c * n, w.dF.data() + c * n, b.ny, b.nx))) return st;
  }
  if (upwind) {
    // upwind_dissipation, lib.rs:619-674 (tmp0/tmp1 reuse ehat/fhat); the
    // first failing point in point order decides the status, as the
    // sequential loop's first assert does
    std::vector<std::pair<size_t, int>> fails;
    std::mutex fmu;
    pfor(n, [&](size_t q0, size_t q1) {
    for (size_t q = q0; q < q1; ++q) {
      double r = rho[q], ru = rhou[q], rv = rhov[q], ee = e[q];
      if (!(r > 0.0)) {
        std::lock_guard<std::mutex> lk(fmu);
        fails.push_back({q, ST_RHO});
        return;
      }
      double u = ru / r, v = rv / r;
      double uhat = b.jxx[q] * u + b.jxy[q] * v;
      double vhat = b.jyx[q] * u + b.jyy[q] * v;
      double pp = pressure(gamma, r, ru, rv, ee);
      if (!(pp > 0.0)) {
        std::lock_guard<std::mutex> lk(fmu);
        fails.push_back({q, ST_PRESSURE});
        return;
      }
      double c = std::sqrt(gamma * pp / r);
      double au = std::fabs(uhat) + c * std::sqrt(b.jxx[q] * b.jxx[q] + b.jxy[q] * b.jxy[q]);
      double av = std::fabs(vhat) + c * std::sqrt(b.jyx[q] * b.jyx[q] + b.jyy[q] * b.jyy[q]);
      eh[q] = au * r;
      fh[q] = av * r;
      eh[n + q] = au * ru;
      fh[n + q] = av * ru;
      eh[2 * n + q] = au * rv;
      fh[2 * n + q] = av * rv;
      eh[3 * n + q] = au * ee;
      fh[3 * n + q] = av * ee;
    }
    });
    if (!fails.empty()) return std::min_element(fails.begin(), fails.end())->second;
    for (int c = 0; c < 4; ++c) {
      if ((st = diff_2d(b.op_xi, KIND_DISS, 0, eh + c * n, w.adx.data() + c * n, b.ny, b.nx))) return st;
      if ((st = diff_2d(b.op_eta, KIND_DISS, 1, fh + c * n, w.ady.data() + c * n, b.ny, b.nx))) return st;
    }
    pfor(n, [&](size_t q0, size_t q1) {
      for (int c = 0; c < 4; ++c)
        for (size_t q = q0; q < q1; ++q)
          kout[c * n + q] = (-w.dE[c * n + q] - w.dF[c * n + q] + w.adx[c * n + q] + w.ady[c * n + q]) / b.detj[q];
    });
  } else {
    pfor(n, [&](size_t q0, size_t q1) {
      for (int c = 0; c < 4; ++c)
        for (size_t q = q0; q < q1; ++q) kout[c * n + q] = (-w.dE[c * n + q] - w.dF[c * n + q]) / b.detj[q];
    });
  }
  if (!S.sat) return ST_OK;  // RHS_no_SAT, lib.rs:518-568
  // SAT_characteristics, lib.rs:1002-1013 (north, south, east, west)
  const OpDef* oe = get_op(b.op_eta);
  const OpDef* ox = get_op(b.op_xi);
  const double hi_eta = (oe->h2 ? (double)(b.ny - 2) : (double)(b.ny - 1)) / oe->h[0];
  const double hi_xi = (ox->h2 ? (double)(b.nx - 2) : (double)(b.nx - 1)) / ox->h[0];
  EdgeView g[4];
  for (int d = 0; d < 4; ++d)
    if ((st = extract_edge(S, bi, fields, d, time, g[d]))) return st;
  const size_t nxs = (size_t)b.nx;
  size_t rN = (size_t)(b.ny - 1) * nxs;
  // north: row ny-1, sign -1, tau +1, metrics J, J·ηx, J·ηy
  sat_characteristic(kout + rN, n, 1, y + rN, n, 1, g[DIR_NORTH].p, g[DIR_NORTH].cs, g[DIR_NORTH].ps, b.nx, hi_eta, -1.0,
                     1.0, &b.detj[rN], &b.jyx[rN], &b.jyy[rN], 1);
  // south: row 0, sign +1, tau -1
  sat_characteristic(kout, n, 1, y, n, 1, g[DIR_SOUTH].p, g[DIR_SOUTH].cs, g[DIR_SOUTH].ps, b.nx, hi_eta, 1.0, -1.0,
                     &b.detj[0], &b.jyx[0], &b.jyy[0], 1);
  // east: column nx-1, sign -1, tau +1, metrics J, J·ξx, J·ξy
  size_t cE = (size_t)(b.nx - 1);
  sat_characteristic(kout + cE, n, nxs, y + cE, n, nxs, g[DIR_EAST].p, g[DIR_EAST].cs, g[DIR_EAST].ps, b.ny, hi_xi, -1.0,
                     1.0, &b.detj[cE], &b.jxx[cE], &b.jxy[cE], nxs);
  // west: column 0, sign +1, tau -1
  sat_characteristic(kout, n, nxs, y, n, nxs, g[DIR_WEST].p, g[DIR_WEST].cs, g[DIR_WEST].ps, b.ny, hi_xi, 1.0, -1.0,
                     &b.detj[0], &b.jxx[0], &b.jxy[0], nxs);
  return ST_OK;
}

bool use_upwind(const System& S, const Block& b, int* err) {
  *err = ST_OK;
  if (S.rhs_kind == RHS_TRAD) return false;
  if (S.rhs_kind == RHS_UPWIND) {
    if (!block_is_upwind(b)) *err = ST_NOT_UPWIND;
    return true;
  }
  return block_is_upwind(b);  // multigrid/src/system.rs:521-525
}

int system_rhs(System& S, const std::vector<std::vector<double>>& fields, double time, std::vector<std::vector<double>>& kout,
               int nthreads) {
  size_t nb = S.blocks.size();
  std::vector<int> status(nb, ST_OK);
  auto run = [&](size_t bi) {
    int err;
    bool up = use_upwind(S, S.blocks[bi], &err);
    if (err) {
      status[bi] = err;
      return;
    }
    status[bi] = block_rhs(S, bi, fields, time, kout[bi].data(), up);
  };
  if (nthreads <= 1 || nb == 1) {
    for (size_t bi = 0; bi < nb; ++bi) run(bi);
  } else {
    // one thread per block, as multigrid's create_distributed (system.rs:160-366);
    // each stage reads only the stage-input fields, so a join is the barrier
    for (size_t b0 = 0; b0 < nb; b0 += (size_t)nthreads) {
      std::vector<std::thread> th;
      for (size_t bi = b0; bi < nb && bi < b0 + (size_t)nthreads; ++bi) th.emplace_back(run, bi);
      for (auto& t : th) t.join();
    }
  }
  for (int s : status)
    if (s) return s;
  return ST_OK;
}

// integrate::<Rk4> (integrate/src/lib.rs:176-222): stage inputs fut = prev +
// (a·dt)·k_j (zero a skipped), final fut = prev + Σ (b_j·dt)·k_j in order; each
// scaled_add is y = y + alpha*x (ndarray).
int system_step(System& S, double dt, int nthreads) {
  static const double A[3][3] = {{0.5, 0, 0}, {0.0, 0.5, 0}, {0.0, 0.0, 1.0}};
  static const double Bw[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  static const double C[3] = {0.5, 0.5, 1.0};
  size_t nb = S.blocks.size();
  auto copy_prev = [&]() {
    for (size_t bi = 0; bi < nb; ++bi) S.fnext[bi] = S.fnow[bi];
  };
  auto scaled_add = [&](int j, double scale) {
    for (size_t bi = 0; bi < nb; ++bi) {
      double* f = S.fnext[bi].data();
      const double* kk = S.k[j][bi].data();
      pfor(S.fnext[bi].size(), [&](size_t q0, size_t q1) {
        for (size_t q = q0; q < q1; ++q) f[q] = f[q] + scale * kk[q];
      });
    }
  };
  int st;
  for (int i = 0; i < 4; ++i) {
    double simtime;
    copy_prev();
    if (i == 0) {
      simtime = S.time;
    } else {
      for (int j = 0; j < i; ++j) {
        double a = A[i - 1][j];
        if (a == 0.0) continue;
        scaled_add(j, a * dt);
      }
      simtime = S.time + dt * C[i - 1];
    }
    if ((st = system_rhs(S, S.fnext, simtime, S.k[i], nthreads))) return st;
  }
  copy_prev();
  for (int j = 0; j < 4; ++j) scaled_add(j, Bw[j] * dt);
  S.time += dt;
  std::swap(S.fnow, S.fnext);
  return ST_OK;
}

// integrate_embedded_rk::<BogackiShampine> (integrate/src/lib.rs:232-251)
// over integrate (:176-222): stage i > 0 takes fut = prev + Σ (a·dt)·k_j and
// time t + dt·C[i-1], where BogackiShampine's C (:150) starts with its 0, so
// the stage times are t, t, t + dt/2, t + 3dt/4; fut = prev + Σ (b·dt)·k_j,
// time += dt; fut2 = prev + Σ (b*·dt)·k_j. fut lands in fnext.
int system_step_bs(System& S, double dt, double& time, std::vector<std::vector<double>>& fut2, int nthreads) {
  static const double A[3][3] = {{1.0 / 2.0, 0.0, 0.0}, {0.0, 3.0 / 4.0, 0.0}, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}};
  static const double Bw[4] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0};
  static const double Bs[4] = {7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0};
  static const double C[4] = {0.0, 1.0 / 2.0, 3.0 / 4.0, 1.0};
  size_t nb = S.blocks.size();
  auto scaled_add = [&](std::vector<std::vector<double>>& f, int j, double scale) {
    for (size_t bi = 0; bi < nb; ++bi) {
      double* y = f[bi].data();
      const double* kk = S.k[j][bi].data();
      size_t m = f[bi].size();
      for (size_t q = 0; q < m; ++q) y[q] = y[q] + scale * kk[q];
    }
  };
  int st;
  for (int i = 0; i < 4; ++i) {
    double simtime = time;
    S.fnext = S.fnow;
    if (i > 0) {
      for (int j = 0; j < i; ++j) {
        if (A[i - 1][j] == 0.0) continue;
        scaled_add(S.fnext, j, A[i - 1][j] * dt);
      }
      simtime = time + dt * C[i - 1];
    }
    if ((st = system_rhs(S, S.fnext, simtime, S.k[i], nthreads))) return st;
  }
  S.fnext = S.fnow;
  for (int j = 0; j < 4; ++j)
    if (Bw[j] != 0.0) scaled_add(S.fnext, j, Bw[j] * dt);
  time += dt;
  fut2 = S.fnow;
  for (int j = 0; j < 4; ++j)
    if (Bs[j] != 0.0) scaled_add(fut2, j, Bs[j] * dt);
  return ST_OK;
}

// integrate::<B> for an explicit tableau given as arrays (integrate/src/lib.rs:
// 17-24, 176-222): ns stages, A row-major ns x ns (stage i > 0 uses A[i][j],
// j < i, the reference's A[i-1][j]), B (ns), C (ns - 1 used: stage i > 0 at
// time t + dt·C[i-1]); zero coefficients are skipped, as the reference does.
int system_step_tableau(System& S, int ns, const double* A, const double* B, const double* C, double dt, int nthreads) {
  size_t nb = S.blocks.size();
  std::vector<std::vector<std::vector<double>>> k(ns, S.fnow);
  auto scaled_add = [&](std::vector<std::vector<double>>& f, int j, double scale) {
    for (size_t bi = 0; bi < nb; ++bi) {
      double* y = f[bi].data();
      const double* kk = k[j][bi].data();
      size_t m = f[bi].size();
      for (size_t q = 0; q < m; ++q) y[q] = y[q] + scale * kk[q];
    }
  };
  int st;
  for (int i = 0; i < ns; ++i) {
    double simtime = S.time;
    S.fnext = S.fnow;
    if (i > 0) {
      for (int j = 0; j < i; ++j) {
        const double a = A[i * ns + j];
        if (a == 0.0) continue;
        scaled_add(S.fnext, j, a * dt);
      }
      simtime = S.time + dt * C[i - 1];
    }
    if ((st = system_rhs(S, S.fnext, simtime, k[i], nthreads))) return st;
  }
  S.fnext = S.fnow;
  for (int j = 0; j < ns; ++j)
    if (B[j] != 0.0) scaled_add(S.fnext, j, B[j] * dt);
  S.time += dt;
  std::swap(S.fnow, S.fnext);
  return ST_OK;
}

// Field::h2_err, lib.rs:377-432
double h2_err_impl(const OpDef* oe, const OpDef* ox, const double* a, const double* b, int ny, int nx) {
  std::vector<double> hx(nx), hy(ny);
  double fx = ox->h2 ? 1.0 / (double)(nx - 2) : 1.0 / (double)(nx - 1);
  double fy = 1.0 / (oe->h2 ? (double)(ny - 2) : (double)(ny - 1));
  for (int i = 0; i < nx; ++i) {
    double h = 1.0;
    if (i < ox->M) h = ox->h[i];
    else if (i >= nx - ox->M) h = ox->h[nx - 1 - i];
    hx[i] = h * fx;
  }
  for (int j = 0; j < ny; ++j) {
    double h = 1.0;
    if (j < oe->M) h = oe->h[j];
    else if (j >= ny - oe->M) h = oe->h[ny - 1 - j];
    hy[j] = h * fy;
  }
  double s = 0.0;
  size_t n = (size_t)nx * ny;
  for (int c = 0; c < 4; ++c)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t q = c * n + (size_t)j * nx + i;
        double d = a[q] - b[q];
        s += d * d * hx[i] * hy[j];
      }
  return std::sqrt(s);
}

}  // namespace

// ------------------------------------------------------------------ C API
extern "C" {

int orc_op_info(int op, int* M, int* N, int* D, int* is_h2, int* is_upwind) {
  const OpDef* d = get_op(op);
  if (!d) return ST_INVALID;
  *M = d->M;
  *N = d->N;
  *D = d->D;
  *is_h2 = d->h2;
  *is_upwind = d->upwind;
  return ST_OK;
}

int orc_op_h(int op, double* h) {
  const OpDef* d = get_op(op);
  if (!d) return ST_INVALID;
  for (int i = 0; i < d->M; ++i) h[i] = d->h[i];
  return ST_OK;
}

int orc_op_block(int op, int kind, double* start, double* diag, double* end) {
  const BlockMat* B = get_block(op, kind);
  if (!B) return get_op(op) ? ST_NOT_UPWIND : ST_INVALID;
  std::memcpy(start, B->start.data(), sizeof(double) * B->M * B->N);
  std::memcpy(diag, B->diag.data(), sizeof(double) * B->D);
  std::memcpy(end, B->end.data(), sizeof(double) * B->M * B->N);
  return ST_OK;
}

int orc_diff(int op, int kind, int axis, const double* prev, double* fut, int ny, int nx) {
  return diff_2d(op, kind, axis, prev, fut, ny, nx);
}

void orc_set_gamma(double g) {
  g_gamma = g;
  g_gamma_set = true;
}

int orc_vortex(const orc_vortex_params* vp, double t, const double* x, const double* y, long n, double* out4) {
  return vortex_eval(*vp, t, x, y, (size_t)n, 1, 1, out4, out4 + n, out4 + 2 * n, out4 + 3 * n, 1);
}

int orc_metrics(int op_eta, int op_xi, const double* x, const double* y, int ny, int nx, double* out5) {
  Block b;
  b.ny = ny;
  b.nx = nx;
  b.op_eta = op_eta;
  b.op_xi = op_xi;
  b.x.assign(x, x + (size_t)nx * ny);
  b.y.assign(y, y + (size_t)nx * ny);
  int st = compute_metrics(b);
  if (st) return st;
  size_t n = b.npts();
  std::memcpy(out5, b.detj.data(), n * 8);
  std::memcpy(out5 + n, b.jxx.data(), n * 8);
  std::memcpy(out5 + 2 * n, b.jxy.data(), n * 8);
  std::memcpy(out5 + 3 * n, b.jyx.data(), n * 8);
  std::memcpy(out5 + 4 * n, b.jyy.data(), n * 8);
  return ST_OK;
}

int orc_h2_err(int op_eta, int op_xi, const double* a, const double* b, int ny, int nx, double* out) {
  const OpDef* oe = get_op(op_eta);
  const OpDef* ox = get_op(op_xi);
  if (!oe || !ox) return ST_INVALID;
  *out = h2_err_impl(oe, ox, a, b, ny, nx);
  return ST_OK;
}

void* orc_system_create(int nblocks, const int* ny, const int* nx, const int* op_eta, const int* op_xi, const double* const* x,
                        const double* const* y, int* status) {
  System* S = new System();
  S->blocks.resize(nblocks);
  S->work.resize(nblocks);
  for (int i = 0; i < nblocks; ++i) {
    Block& b = S->blocks[i];
    b.ny = ny[i];
    b.nx = nx[i];
    b.op_eta = op_eta[i];
    b.op_xi = op_xi[i];
    if (!get_op(b.op_eta) || !get_op(b.op_xi)) {
      *status = ST_INVALID;
      delete S;
      return nullptr;
    }
    size_t n = b.npts();
    b.x.assign(x[i], x[i] + n);
    b.y.assign(y[i], y[i] + n);
    int st = compute_metrics(b);
    if (st) {
      *status = st;
      delete S;
      return nullptr;
    }
    S->work[i].ensure(n, b.nx, b.ny);
    std::memset(b.bc_vortex, 0, sizeof(b.bc_vortex));
  }
  S->fnow.resize(nblocks);
  S->fnext.resize(nblocks);
  for (int j = 0; j < 4; ++j) S->k[j].resize(nblocks);
  for (int i = 0; i < nblocks; ++i) {
    size_t n = S->blocks[i].npts();
    S->fnow[i].assign(4 * n, 0.0);
    S->fnext[i].assign(4 * n, 0.0);
    for (int j = 0; j < 4; ++j) S->k[j][i].assign(4 * n, 0.0);
  }
  *status = ST_OK;
  return S;
}

void orc_system_destroy(void* s) { delete (System*)s; }

int orc_system_set_bc(void* s, int block, int dir, int kind, int neighbour, const orc_vortex_params* vp) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size() || dir < 0 || dir > 3) return ST_INVALID;
  Block& b = S->blocks[block];
  b.bc_kind[dir] = kind;
  b.bc_nb[dir] = neighbour;
  if (vp) b.bc_vortex[dir] = *vp;
  return ST_OK;
}

// BoundaryCharacteristic::Eval(evaluator) (lib.rs:741) through a callback
int orc_system_set_bc_eval(void* s, int block, int dir, orc_eval_fn fn, void* user) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size() || dir < 0 || dir > 3 || !fn) return ST_INVALID;
  Block& b = S->blocks[block];
  b.bc_kind[dir] = BC_EVAL;
  b.bc_eval[dir] = fn;
  b.bc_eval_user[dir] = user;
  return ST_OK;
}

// BoundaryCharacteristic::Interpolate(neighbour, operator) (lib.rs:743)
int orc_system_set_bc_interpolate(void* s, int block, int dir, int neighbour, int iop) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size() || dir < 0 || dir > 3 || iop < 0 || iop >= IOP_COUNT) return ST_INVALID;
  Block& b = S->blocks[block];
  b.bc_kind[dir] = BC_INTERPOLATE;
  b.bc_nb[dir] = neighbour;
  b.bc_iop[dir] = iop;
  return ST_OK;
}

// BoundaryCharacteristic::MultiGrid(pieces) (lib.rs:744): npieces triples
// (block, start, end)
int orc_system_set_bc_multigrid(void* s, int block, int dir, int npieces, const int* triples) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size() || dir < 0 || dir > 3 || npieces < 1 || !triples) return ST_INVALID;
  Block& b = S->blocks[block];
  b.bc_kind[dir] = BC_MULTIGRID;
  b.bc_pieces[dir].assign(triples, triples + 3 * (size_t)npieces);
  return ST_OK;
}

// Boundary data g of one edge of one block for the current state at time t,
// as boundary_extract assembles it (lib.rs:804-867): 4 x (edge length).
int orc_system_edge(void* s, int block, int dir, double t, double* out) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size() || dir < 0 || dir > 3) return ST_INVALID;
  const Block& b = S->blocks[block];
  S->work[block].ensure(b.npts(), b.nx, b.ny);
  EdgeView v;
  const int st = extract_edge(*S, (size_t)block, S->fnow, dir, t, v);
  if (st) return st;
  const int n = (dir == DIR_NORTH || dir == DIR_SOUTH) ? b.nx : b.ny;
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < n; ++i) out[(size_t)c * n + i] = v.p[c * v.cs + (size_t)i * v.ps];
  return ST_OK;
}

// InterpolationOperator::fine2coarse / coarse2fine (sbp/src/operators.rs:108-113)
int orc_interpolate(int iop, int f2c, const double* in, int nin, double* out, int nout) {
  InterpDef d;
  if (!interp_def(iop, f2c != 0, d)) return ST_INVALID;
  if (!interp_lengths(d, f2c != 0, nin, nout)) return ST_SHAPE;
  interpolate(d, in, 1, nin, out, 1, nout);
  return ST_OK;
}

// Table shapes and coefficients of one direction of an operator.
int orc_interp_info(int iop, int f2c, int* shape6) {
  InterpDef d;
  if (!interp_def(iop, f2c != 0, d)) return ST_INVALID;
  const int v[6] = {d.brows, d.bcols, d.drows, d.dcols, d.j0, d.j1};
  std::memcpy(shape6, v, sizeof v);
  return ST_OK;
}

int orc_interp_coefficients(int iop, int f2c, double* block, double* diag) {
  InterpDef d;
  if (!interp_def(iop, f2c != 0, d)) return ST_INVALID;
  std::memcpy(block, d.block, sizeof(double) * d.brows * d.bcols);
  std::memcpy(diag, d.diag, sizeof(double) * d.drows * d.dcols);
  return ST_OK;
}

int orc_system_set_rhs_kind(void* s, int kind) {
  ((System*)s)->rhs_kind = kind;
  return ST_OK;
}

int orc_system_metrics(void* s, int block, double* out5) {
  System* S = (System*)s;
  Block& b = S->blocks[block];
  size_t n = b.npts();
  std::memcpy(out5, b.detj.data(), n * 8);
  std::memcpy(out5 + n, b.jxx.data(), n * 8);
  std::memcpy(out5 + 2 * n, b.jxy.data(), n * 8);
  std::memcpy(out5 + 3 * n, b.jyx.data(), n * 8);
  std::memcpy(out5 + 4 * n, b.jyy.data(), n * 8);
  return ST_OK;
}

int orc_system_set_state(void* s, int block, const double* state) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size()) return ST_INVALID;
  std::memcpy(S->fnow[block].data(), state, S->fnow[block].size() * 8);
  return ST_OK;
}

int orc_system_get_state(void* s, int block, double* state) {
  System* S = (System*)s;
  if (block < 0 || block >= (int)S->blocks.size()) return ST_INVALID;
  std::memcpy(state, S->fnow[block].data(), S->fnow[block].size() * 8);
  return ST_OK;
}

int orc_system_vortex(void* s, double t, const orc_vortex_params* vp) {
  System* S = (System*)s;
  for (size_t i = 0; i < S->blocks.size(); ++i) {
    Block& b = S->blocks[i];
    size_t n = b.npts();
    double* f = S->fnow[i].data();
    int st = vortex_eval(*vp, t, b.x.data(), b.y.data(), n, 1, 1, f, f + n, f + 2 * n, f + 3 * n, 1);
    if (st) return st;
  }
  return ST_OK;
}

// RHS of the current state at time t into k (concatenated per block).
int orc_system_rhs(void* s, double t, double* kout) {
  System* S = (System*)s;
  std::vector<std::vector<double>> k(S->blocks.size());
  for (size_t i = 0; i < S->blocks.size(); ++i) k[i].assign(S->fnow[i].size(), 0.0);
  int st = system_rhs(*S, S->fnow, t, k, 1);
  if (st) return st;
  size_t off = 0;
  for (auto& kk : k) {
    std::memcpy(kout + off, kk.data(), kk.size() * 8);
    off += kk.size();
  }
  return ST_OK;
}

// Threads used inside each block (1 = the reference's one thread per block;
// the results do not depend on it)
int orc_set_inner_threads(int n) {
  if (n < 1) return ST_INVALID;
  g_inner_threads = n;
  return ST_OK;
}

// RHS_no_SAT (lib.rs:518-568) for on = 0, the full RHS otherwise
int orc_system_set_sat(void* s, int on) {
  ((System*)s)->sat = on != 0;
  return ST_OK;
}

// One SAT function (SAT_north / south / east / west, lib.rs:1025-1144) on
// caller arrays: k, y (4, ny, nx), metrics (5, ny, nx), g (4, n); k += SAT.
int orc_sat_edge(int op_eta, int op_xi, int dir, int ny, int nx, double* k, const double* y, const double* met, const double* g) {
  const OpDef* oe = get_op(op_eta);
  const OpDef* ox = get_op(op_xi);
  if (!oe || !ox || dir < 0 || dir > 3) return ST_INVALID;
  if (!g_gamma_set) return ST_GAMMA;
  const size_t n = (size_t)nx * ny, nxs = (size_t)nx;
  const double hi_eta = (oe->h2 ? (double)(ny - 2) : (double)(ny - 1)) / oe->h[0];
  const double hi_xi = (ox->h2 ? (double)(nx - 2) : (double)(nx - 1)) / ox->h[0];
  const double *J = met, *jxx = met + n, *jxy = met + 2 * n, *jyx = met + 3 * n, *jyy = met + 4 * n;
  const size_t rN = (size_t)(ny - 1) * nxs, cE = nxs - 1;
  switch (dir) {
    case DIR_NORTH:
      sat_characteristic(k + rN, n, 1, y + rN, n, 1, g, (size_t)nx, 1, nx, hi_eta, -1.0, 1.0, J + rN, jyx + rN, jyy + rN, 1);
      break;
    case DIR_SOUTH:
      sat_characteristic(k, n, 1, y, n, 1, g, (size_t)nx, 1, nx, hi_eta, 1.0, -1.0, J, jyx, jyy, 1);
      break;
    case DIR_EAST:
      sat_characteristic(k + cE, n, nxs, y + cE, n, nxs, g, (size_t)ny, 1, ny, hi_xi, -1.0, 1.0, J + cE, jxx + cE, jxy + cE,
                         nxs);
      break;
    default:
      sat_characteristic(k, n, nxs, y, n, nxs, g, (size_t)ny, 1, ny, hi_xi, 1.0, -1.0, J, jxx, jxy, nxs);
  }
  return ST_OK;
}

// nsteps of integrate::<B> for the tableau (ns, A, B, C)
int orc_system_integrate(void* s, int ns, const double* A, const double* B, const double* C, double dt, long nsteps,
                         int nthreads) {
  System* S = (System*)s;
  if (ns < 1) return ST_INVALID;
  for (long i = 0; i < nsteps; ++i) {
    int st = system_step_tableau(*S, ns, A, B, C, dt, nthreads);
    if (st) return st;
  }
  return ST_OK;
}

int orc_system_advance(void* s, double dt, long nsteps, int nthreads) {
  System* S = (System*)s;
  for (long i = 0; i < nsteps; ++i) {
    int st = system_step(*S, dt, nthreads);
    if (st) return st;
  }
  return ST_OK;
}

// One embedded step at loop time *time (advanced by dt): the state becomes
// fut; fut2 (all blocks concatenated) is written to out2.
int orc_system_embedded_step(void* s, double* time, double dt, double* out2, int nthreads) {
  System* S = (System*)s;
  std::vector<std::vector<double>> fut2;
  int st = system_step_bs(*S, dt, *time, fut2, nthreads);
  if (st) return st;
  size_t off = 0;
  for (auto& f : fut2) {
    std::memcpy(out2 + off, f.data(), f.size() * 8);
    off += f.size();
  }
  std::swap(S->fnow, S->fnext);
  return ST_OK;
}

// euler::System::advance_adaptive (lib.rs:149-185), for every block of the
// system with its own boundary settings: the loop time starts at 0 and
// grows by the trial step inside integrate and again on acceptance; the
// error is h2_err(previous state, fut2), summed over blocks; accepted: swap,
// guess ·1.05; rejected: guess ·0.8.
int orc_system_advance_adaptive(void* s, double dt, double* guess, double maxerr, int nthreads) {
  System* S = (System*)s;
  double time = 0.0;
  std::vector<std::vector<double>> fut2;
  while (time < dt) {
    int st = system_step_bs(*S, *guess, time, fut2, nthreads);
    if (st) return st;
    double err = 0.0;
    for (size_t i = 0; i < S->blocks.size(); ++i) {
      const Block& b = S->blocks[i];
      err += h2_err_impl(get_op(b.op_eta), get_op(b.op_xi), S->fnow[i].data(), fut2[i].data(), b.ny, b.nx);
    }
    if (err < maxerr) {
      time += *guess;
      std::swap(S->fnow, S->fnext);
      *guess *= 1.05;
    } else {
      *guess *= 0.8;
    }
  }
  return ST_OK;
}

double orc_system_time(void* s) { return ((System*)s)->time; }
void orc_system_set_time(void* s, double t) { ((System*)s)->time = t; }

// SingleThreadedSystem::max_dt, system.rs:542-589
int orc_system_max_dt(void* s, double* out) {
  System* S = (System*)s;
  bool is_h2 = false;
  for (auto& b : S->blocks) is_h2 = is_h2 || get_op(b.op_xi)->h2 || get_op(b.op_eta)->h2;
  double c_max = is_h2 ? 0.5 : 1.0;
  double max_dt = INFINITY;
  for (size_t i = 0; i < S->blocks.size(); ++i) {
    Block& b = S->blocks[i];
    size_t n = b.npts();
    const double* f = S->fnow[i].data();
    double mu = 0.0, mv = 0.0;
    for (size_t q = 0; q < n; ++q) {
      double u = f[n + q] / f[q], v = f[2 * n + q] / f[q];
      double uhat = b.jxx[q] * u + b.jxy[q] * v;
      double vhat = b.jyx[q] * u + b.jyy[q] * v;
      mu = std::fmax(mu, std::fabs(uhat));
      mv = std::fmax(mv, std::fabs(vhat));
    }
    double dx = 1.0 / (double)b.nx, dy = 1.0 / (double)b.ny;
    double c_dt = std::fmax(mu / dx, mv / dy);
    max_dt = std::fmin(max_dt, c_max / c_dt);
  }
  *out = max_dt;
  return ST_OK;
}

// System::error, system.rs:618-635: Σ_blocks h2_err(q, vortex(t))
int orc_system_error_vortex(void* s, const orc_vortex_params* vp, double* out) {
  System* S = (System*)s;
  double e = 0.0;
  for (size_t i = 0; i < S->blocks.size(); ++i) {
    Block& b = S->blocks[i];
    size_t n = b.npts();
    std::vector<double> fv(4 * n);
    int st = vortex_eval(*vp, S->time, b.x.data(), b.y.data(), n, 1, 1, fv.data(), fv.data() + n, fv.data() + 2 * n,
                         fv.data() + 3 * n, 1);
    if (st) return st;
    e += h2_err_impl(get_op(b.op_eta), get_op(b.op_xi), S->fnow[i].data(), fv.data(), b.ny, b.nx);
  }
  *out = e;
  return ST_OK;
}

}  // extern "C"
