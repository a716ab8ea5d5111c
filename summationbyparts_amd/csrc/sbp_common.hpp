// Device helpers shared by the HIP kernels: coefficient access, the Euler
// pointwise flux/dissipation terms, the characteristic SAT and the vortex.
#pragma once
#include <hip/hip_runtime.h>

#include "sbp_device.hpp"
#include "sbp_operators.hpp"

namespace sbp {

// Reciprocal and square root from the hardware estimates (v_rcp_f64 /
// v_rsq_f64, ~2^-23 relative) refined by Newton-Raphson / Goldschmidt steps
// to within about one ulp, without the range scaling of the IEEE-exact
// library versions. Arguments are normal, positive numbers (densities,
// squared sound speeds); checks on them happen before use.
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Load through a uniform base pointer and a 32-bit element offset (selects
// the SGPR-base + VGPR-offset addressing of global loads).
template <typename T>
__device__ __forceinline__ T ldg(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (idx << (sizeof(T) == 8 ? 3 : 2)));
}
template <typename T>
__device__ __forceinline__ void stg(T* base, uint32_t idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (idx << (sizeof(T) == 8 ? 3 : 2))) = v;
}
// Streaming store/load (non-temporal hint): for data written or read once per
// stage, so that it neither allocates nor evicts L2 lines holding halos.
template <typename T>
__device__ __forceinline__ void stg_nt(T* base, uint32_t idx, T v) {
  using GP = __attribute__((address_space(1))) T*;
  using GC = __attribute__((address_space(1))) char*;
  __builtin_nontemporal_store(v, (GP)((GC)base + (idx << (sizeof(T) == 8 ? 3 : 2))));
}
template <typename T>
__device__ __forceinline__ T ldg_nt(const T* base, uint32_t idx) {
  using GP = const __attribute__((address_space(1))) T*;
  using GC = const __attribute__((address_space(1))) char*;
  return __builtin_nontemporal_load((GP)((GC)base + (idx << (sizeof(T) == 8 ? 3 : 2))));
}

// The sticky device error word keeps the first error raised (the reference
// panics at its first failed assertion): a thread with several failures at
// one point reports the lowest bit, the check the reference makes first
// (density before pressure).
__device__ __forceinline__ void record_error(unsigned* word, unsigned bits) {
  atomicCAS(word, 0u, bits & (~bits + 1u));
}

// Error bits of the sticky device error word.
constexpr unsigned ERRBIT_RHO = 1u << SBP_ERR_RHO;
constexpr unsigned ERRBIT_PRESSURE = 1u << SBP_ERR_PRESSURE;
constexpr unsigned ERRBIT_DETJ = 1u << SBP_ERR_DETJ;

// Coefficients of one operator block (BlockMatrix, algos.rs:31-41): interior
// diagonal, start block and end block. End block entries are start-block
// entries read flipped in both directions; DIFF additionally flips the sign.
template <class OP, int KIND>
struct Blk;

template <class OP>
struct Blk<OP, SBP_DIFF> {
  static constexpr int M = OP::M, N = OP::N, D = OP::D, H = OP::H;
  __device__ __forceinline__ static constexpr double diag(int d) { return OP::diff_diag[d]; }
  __device__ __forceinline__ static double start(int r, int k) { return OP::diff_start[r][k]; }
  __device__ __forceinline__ static double end(int r, int k) { return -OP::diff_start[M - 1 - r][N - 1 - k]; }
};

template <class OP>
struct Blk<OP, SBP_DISS> {
  static constexpr int M = OP::M, N = OP::N, D = OP::D, H = OP::H;
  __device__ __forceinline__ static constexpr double diag(int d) { return OP::diss_diag[d]; }
  __device__ __forceinline__ static double start(int r, int k) { return OP::diss_start[r][k]; }
  __device__ __forceinline__ static double end(int r, int k) { return OP::diss_start[M - 1 - r][N - 1 - k]; }
};

// Interior stencil Σ_d diag[d] p[(d-H)·stride], p pointing at the output row.
template <class B, typename T>
__device__ __forceinline__ T interior_sum(const T* p, int stride) {
  T s = T(0);
#pragma unroll
  for (int d = 0; d < B::D; ++d) {
    if (B::diag(d) != 0.0) s += T(B::diag(d)) * p[(d - B::H) * stride];
  }
  return s;
}

// Full operator row i of an axis of length n: start block, interior or end
// block (algos.rs:101-187). `line` points at axis index r0, stride `stride`.
// The closure loops are kept rolled: their coefficient rows are selected per
// lane at run time, and unrolling them would hold a whole row in registers.
template <class B, typename T>
__device__ __forceinline__ T row_sum(const T* line, int stride, int i, int n, int r0) {
  if (i < B::M) {
    T s = T(0);
#pragma unroll 1
    for (int k = 0; k < B::N; ++k) s += T(B::start(i, k)) * line[(k - r0) * stride];
    return s;
  }
  if (i >= n - B::M) {
    const int r = i - (n - B::M);
    const int b0 = n - B::N - r0;
    T s = T(0);
#pragma unroll 1
    for (int k = 0; k < B::N; ++k) s += T(B::end(r, k)) * line[(b0 + k) * stride];
    return s;
  }
  return interior_sum<B>(line + (i - r0) * stride, stride);
}

// Tile placement along one axis of length n with K = ceil(n/T) tiles: tile t
// computes [x0, x0+T) and writes [w0, w1). The last tile is right-aligned and
// owns every end-closure row, so a halo of H points always covers the rows a
// tile writes (closure rows read up to N-1 points from the edge).
__device__ __forceinline__ void tile_place(int t, int K, int n, int T, int M, int& x0, int& w0, int& w1) {
  if (K <= 1) {
    x0 = 0;
    w0 = 0;
    w1 = n;
    return;
  }
  const int b = min((K - 1) * T, n - M);
  if (t < K - 1) {
    x0 = t * T;
    w0 = t * T;
    w1 = min((t + 1) * T, b);
  } else {
    x0 = n - T;
    w0 = b;
    w1 = n;
  }
}

// Euler flux through a coordinate face with metric pair (jx, jy) and the
// upwind dissipation weight times the state (euler/src/lib.rs:619-735):
//   û = jx u + jy v,  flux = [ρû, ρu û + jx p, ρv û + jy p, (e+p) û]
//   α = |û| + c √(jx²+jy²),  aq = α q.
//   c √(jx²+jy²) is taken as one square root, √(γ p/ρ · (jx²+jy²)).
template <bool UP>
__device__ __forceinline__ void face_terms(double rho, double ru, double rv, double e, double jx, double jy, double gamma,
                                           double f[4], double aq[4], double& alpha, unsigned& err) {
  const double ir = frcp(rho);
  const double u = ru * ir, v = rv * ir;
  const double p = (gamma - 1.0) * (e - (ru * ru + rv * rv) * (0.5 * ir));
  const double uh = jx * u + jy * v;
  f[0] = rho * uh;
  f[1] = ru * uh + jx * p;
  f[2] = rv * uh + jy * p;
  f[3] = (e + p) * uh;
  if constexpr (UP) {
    if (!(rho > 0.0)) err |= ERRBIT_RHO;
    if (!(p > 0.0)) err |= ERRBIT_PRESSURE;
    const double c2g = gamma * p * ir * (jx * jx + jy * jy);
    const double al = fabs(uh) + fsqrt(fmax(c2g, 1e-300));
    alpha = al;
    aq[0] = al * rho;
    aq[1] = al * ru;
    aq[2] = al * rv;
    aq[3] = al * e;
  }
}

// Characteristic SAT of one edge point (SAT_characteristic, lib.rs:1150-1252):
// k += hi·tau·T((L + sign|L|)∘(TI(q - g))).
__device__ __forceinline__ void sat_point(double k[4], const double q[4], const double g[4], double hi, double sign, double tau,
                                          double detj, double ddx, double ddy, double gamma) {
  const double rho = q[0], rhou = q[1], rhov = q[2], e = q[3];
  const double kx_ = ddx / detj, ky_ = ddy / detj;
  const double r = hypot(kx_, ky_);
  const double kx = kx_ / r, ky = ky_ / r;
  const double u = rhou / rho, v = rhov / rho;
  const double theta = kx * u + ky * v;
  const double p = (gamma - 1.0) * (e - (rhou * rhou + rhov * rhov) / (2.0 * rho));
  const double c = sqrt(gamma * p / rho);
  const double phi2 = (gamma - 1.0) * (u * u + v * v) / 2.0;
  const double alpha = rho / (1.41421356237309504880 * c);
  const double phi2_c2 = (phi2 + c * c) / (gamma - 1.0);
  const double T[4][4] = {
      {1.0, 0.0, alpha, alpha},
      {u, ky, alpha * (u + kx * c), alpha * (u - kx * c)},
      {v, -kx, alpha * (v + ky * c), alpha * (v - ky * c)},
      {phi2 / (gamma - 1.0), rho * (ky * u - kx * v), alpha * (phi2_c2 + c * theta), alpha * (phi2_c2 - c * theta)},
  };
  const double U = kx_ * u + ky_ * v;
  const double ch = c * r;
  const double L[4] = {U, U, U + ch, U - ch};
  const double beta = 1.0 / (2.0 * c * c);
  const double g1 = gamma - 1.0;
  const double ic2 = 1.0 / (c * c);
  const double TI[4][4] = {
      {1.0 - phi2 * ic2, g1 * u * ic2, g1 * v * ic2, -g1 * ic2},
      {-(ky * u - kx * v) / rho, ky / rho, -kx / rho, 0.0},
      {beta * (phi2 - c * theta), beta * (kx * c - g1 * u), beta * (ky * c - g1 * v), beta * g1},
      {beta * (phi2 + c * theta), -beta * (kx * c + g1 * u), -beta * (ky * c + g1 * v), beta * g1},
  };
  double res[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) res[i] = q[i] - g[i];
  double lt[4];
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    double t = 0.0;
#pragma unroll
    for (int col = 0; col < 4; ++col) t += TI[row][col] * res[col];
    lt[row] = (L[row] + sign * fabs(L[row])) * t;
  }
  const double ht = hi * tau;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    double t = 0.0;
#pragma unroll
    for (int col = 0; col < 4; ++col) t += T[row][col] * lt[col];
    k[row] += ht * t;
  }
}

// Isentropic vortex superposition at one point (vortex.rs:23-113).
__device__ __forceinline__ void vortex_point(const VortexDev& vp, double gamma, double time, double x, double y, double q[4],
                                             unsigned& err) {
  const double m = vp.mach;
  const double p_inf = 1.0 / (gamma * m * m);
  const double rho_inf = 1.0, u_inf = 1.0, v_inf = 0.0;
  const double e_inf = p_inf / (gamma - 1.0) + rho_inf * (u_inf * u_inf + v_inf * v_inf) / 2.0;
  const double PI = 3.14159265358979323846;
  if (vp.n <= 0) {
    // vortex.rs:56-60: rhou is written twice (ending at rho_inf·v_inf), rhov untouched
    q[0] = rho_inf;
    q[1] = rho_inf * v_inf;
    q[3] = e_inf;
    return;
  }
  for (int iv = 0; iv < vp.n; ++iv) {
    const double rstar = vp.rstar[iv], eps = vp.eps[iv];
    const double dx = (x - vp.x0[iv]) - time;
    const double dy = y - vp.y0[iv];
    const double f = (1.0 - (dx * dx + dy * dy)) / (rstar * rstar);
    const double rv_ =
        pow(1.0 - eps * eps * (gamma - 1.0) * m * m / (8.0 * PI * PI * p_inf * rstar * rstar) * exp(f), 1.0 / (gamma - 1.0));
    if (!(rv_ > 0.0)) err |= ERRBIT_RHO;
    const double p = pow(rv_, gamma) * p_inf;
    const double den = 2.0 * PI * sqrt(p_inf) * rstar * rstar;
    const double ef = exp(f / 2.0);
    const double u = 1.0 - eps * dy / den * ef;
    const double v = eps * dx / den * ef;
    if (!(p > 0.0)) err |= ERRBIT_PRESSURE;
    if (iv == 0) {
      q[0] = rv_;
      q[1] = rv_ * u;
      q[2] = rv_ * v;
      q[3] = p / (gamma - 1.0) + rv_ * (u * u + v * v) / 2.0;
    } else {
      q[0] += rv_ - rho_inf;
      if (!(q[0] > 0.0)) err |= ERRBIT_RHO;
      q[1] += rv_ * u - rho_inf * u_inf;
      q[2] += rv_ * v - rho_inf * v_inf;
      q[3] += (p / (gamma - 1.0) + rv_ * (u * u + v * v) / 2.0) - e_inf;
    }
  }
}

__device__ __forceinline__ const double* stage_input(const BlockDev& B, int mode) {
  switch (mode) {
    case MODE_S0: return B.buf[BUF_YN];
    case MODE_S1: return B.buf[BUF_YA];
    case MODE_S2: return B.buf[BUF_YB];
    case MODE_S3: return B.buf[BUF_YA];
    default: return B.rhs_in;
  }
}

// Boundary data g of edge `dir` at edge index t (boundary_extracts, lib.rs:869-915).
__device__ __forceinline__ void edge_data(const StageArgs& A, const BlockDev& B, int dir, int t, int x, int y, int mode,
                                          double time, double g[4], unsigned& err) {
  const int src = B.edge_src[dir];
  if (src == EDGE_THIS || src == EDGE_GRID) {
    const BlockDev& S = src == EDGE_THIS ? B : A.blocks[B.edge_nb[dir]];
    const double* in = stage_input(S, mode);
    int64_t o;
    switch (dir) {
      case SBP_NORTH: o = t; break;                                  // source south row
      case SBP_SOUTH: o = (int64_t)(S.ny - 1) * S.ld + t; break;     // source north row
      case SBP_EAST: o = (int64_t)t * S.ld; break;                   // source west column
      default: o = (int64_t)t * S.ld + (S.nx - 1); break;             // source east column
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) g[c] = in[c * S.ps + o];
  } else if (src == EDGE_VORTEX) {
    const int64_t o = (int64_t)y * B.ld + x;
    g[2] = 0.0;
    vortex_point(B.vortex[dir], A.gamma, time, B.xc[o], B.yc[o], g, err);
  } else {
    const double* p = B.edge_ptr[dir];
#pragma unroll
    for (int c = 0; c < 4; ++c) g[c] = p[c * B.edge_cs[dir] + t];
  }
}

}  // namespace sbp
