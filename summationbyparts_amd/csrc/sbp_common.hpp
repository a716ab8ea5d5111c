// Device helpers shared by the HIP kernels: coefficient access, the Euler
// pointwise flux/dissipation terms, the characteristic SAT and the vortex.
#pragma once
#include <hip/hip_runtime.h>

#include "sbp_device.hpp"
#include "sbp_expr.hpp"
#include "sbp_interp.hpp"
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

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Load through a uniform base pointer and a 32-bit element offset (selects
// the SGPR-base + VGPR-offset addressing of global loads). The pointers come
// from device descriptors, so the cast to the global address space is what
// keeps the compiler from emitting flat loads (which also wait on LDS).
using f4 = float __attribute__((ext_vector_type(4)));
using d2 = double __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ constexpr uint32_t elt_shift() {
  return sizeof(T) == 16 ? 4 : sizeof(T) == 8 ? 3 : 2;
}
template <typename T>
__device__ __forceinline__ T ldg(const T* base, uint32_t idx) {
  using GP = const __attribute__((address_space(1))) T*;
  using GC = const __attribute__((address_space(1))) char*;
  return *(GP)((GC)base + (idx << elt_shift<T>()));
}
template <typename T>
__device__ __forceinline__ void stg(T* base, uint32_t idx, T v) {
  using GP = __attribute__((address_space(1))) T*;
  using GC = __attribute__((address_space(1))) char*;
  *(GP)((GC)base + (idx << elt_shift<T>())) = v;
}
// Streaming store/load (non-temporal hint): for data written or read once per
// stage, so that it neither allocates nor evicts L2 lines holding halos.
template <typename T>
__device__ __forceinline__ void stg_nt(T* base, uint32_t idx, T v) {
  using GP = __attribute__((address_space(1))) T*;
  using GC = __attribute__((address_space(1))) char*;
  __builtin_nontemporal_store(v, (GP)((GC)base + (idx << elt_shift<T>())));
}
template <typename T>
__device__ __forceinline__ T ldg_nt(const T* base, uint32_t idx) {
  using GP = const __attribute__((address_space(1))) T*;
  using GC = const __attribute__((address_space(1))) char*;
  return __builtin_nontemporal_load((GP)((GC)base + (idx << elt_shift<T>())));
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
constexpr unsigned ERRBIT_EXPR = 1u << SBP_ERR_EXPR;

// Coefficients of one operator block (BlockMatrix, algos.rs:31-41): interior
// diagonal, start block and end block. End block entries are start-block
// entries read flipped in both directions; DIFF additionally flips the sign.
template <class OP, int KIND>
struct Blk;

template <class OP>
struct Blk<OP, SBP_DIFF> {
  static constexpr int M = OP::M, N = OP::N, D = OP::D, H = OP::H;
  static constexpr double end_sign = -1.0;
  static constexpr int table = 0;  // position of its start block in a DIFF, DISS pair of tables
  __device__ __forceinline__ static constexpr double diag(int d) { return OP::diff_diag[d]; }
  __device__ __forceinline__ static double start(int r, int k) { return OP::diff_start[r][k]; }
  __device__ __forceinline__ static double end(int r, int k) { return -OP::diff_start[M - 1 - r][N - 1 - k]; }
};

template <class OP>
struct Blk<OP, SBP_DISS> {
  static constexpr int M = OP::M, N = OP::N, D = OP::D, H = OP::H;
  static constexpr double end_sign = 1.0;
  static constexpr int table = 1;
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
// The coefficient row of a closure is selected per lane at run time (read
// from the constant tables); the N taps are unrolled so that their loads are
// in flight together. CL = false: the caller never asks for a closure row
// (the stage kernel's tiles away from block edges), and the closure code is
// left out.
template <class B, bool CL = true, typename T>
__device__ __forceinline__ T row_sum(const T* line, int stride, int i, int n, int r0) {
  if constexpr (CL) {
    if (i < B::M) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < B::N; ++k) s += T(B::start(i, k)) * line[(k - r0) * stride];
      return s;
    }
    if (i >= n - B::M) {
      const int r = i - (n - B::M);
      const int b0 = n - B::N - r0;
      T s = T(0);
#pragma unroll
      for (int k = 0; k < B::N; ++k) s += T(B::end(r, k)) * line[(b0 + k) * stride];
      return s;
    }
  }
  return interior_sum<B>(line + (i - r0) * stride, stride);
}

// row_sum of the four components at once (component q at line + q·qs): each
// closure coefficient is read from the tables once for all four, the sums
// run in row_sum's order.
template <class B, bool CL = true, typename T>
__device__ __forceinline__ void row_sum4(const T* line, int qs, int stride, int i, int n, int r0, T out[4]) {
  if constexpr (CL) {
    if (i < B::M) {
      T s[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 2
      for (int k = 0; k < B::N; ++k) {
        const T c = T(B::start(i, k));
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += c * line[q * qs + (k - r0) * stride];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) out[q] = s[q];
      return;
    }
    if (i >= n - B::M) {
      const int r = i - (n - B::M);
      const int b0 = n - B::N - r0;
      T s[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 2
      for (int k = 0; k < B::N; ++k) {
        const T c = T(B::end(r, k));
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += c * line[q * qs + (b0 + k) * stride];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) out[q] = s[q];
      return;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = interior_sum<B>(line + q * qs + (i - r0) * stride, stride);
}

// row_sum4 of a closure row (start or end block) with the start block's
// coefficients read from a copy in LDS (`tab`, M x N row-major): a closure
// row's coefficient row differs from lane to lane, and LDS serves such reads
// at a fraction of the constant tables' latency. Interior rows as row_sum4.
// A wavefront holding both closure and interior rows takes the closure path
// for all of them, interior rows with the interior diagonal padded by N - D
// zeros on each side (`diagp`, 2N - D entries in LDS) over an N-wide window
// inside the loaded region [r0, ...): the same taps and products, without
// running both paths in one wavefront.
template <class B, typename T>
__device__ __forceinline__ void row_sum4_tab(const T* line, int qs, int stride, int i, int n, int r0, const double* tab,
                                             const double* diagp, T out[4]) {
  constexpr int PAD = B::N - B::D;
  const bool clos = i < B::M || i >= n - B::M;
  if (__any(clos)) {
    const double* row;
    int dk, b0;
    T sg;
    if (clos) {
      const bool st = i < B::M;
      // end rows: start row M-1-r read backwards, with the kind's sign
      const int r = st ? i : i - (n - B::M);
      row = st ? tab + i * B::N : tab + (B::M - 1 - r) * B::N + (B::N - 1);
      dk = st ? 1 : -1;
      sg = st ? T(1) : T(B::end_sign);
      b0 = st ? -r0 : n - B::N - r0;
    } else {
      // window [w, w + N) holds the stencil [i - H, i + H]; it starts PAD
      // before the stencil unless that leaves the region (the region's end
      // is at least H + 1 past i, and the region is at least N wide)
      const int w = max(i - B::H - PAD, r0);
      row = diagp + (w - (i - B::H - PAD));
      dk = 1;
      sg = T(1);
      b0 = w - r0;
    }
    T s[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int k = 0; k < B::N; ++k) {
      const T c = sg * T(row[dk * k]);
#pragma unroll
      for (int q = 0; q < 4; ++q) s[q] += c * line[q * qs + (b0 + k) * stride];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = s[q];
    return;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = interior_sum<B>(line + q * qs + (i - r0) * stride, stride);
}

// Neumaier (compensated) sums: a running sum and its compensation.
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

// Lane 0 of each wavefront ends with the merge of its 64 lanes.
__device__ __forceinline__ KSum wave_ksum(KSum a) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const KSum b = {__shfl_down(a.s, off, 64), __shfl_down(a.c, off, 64)};
    ksum_merge(a, b);
  }
  return a;
}

// h2_err weight of point (i, j) of an n x m block (lib.rs:389-421): the
// diagonal norms [H, 1.., rev(H)] of both operators, scaled by 1/(n-1)
// (1/(n-2) for h2 operators).
template <class OX, class OY>
__device__ __forceinline__ double hweight(int i, int n, int j, int m) {
  double hx = 1.0, hy = 1.0;
  if (i < OX::M) hx = OX::h[i];
  else if (i >= n - OX::M) hx = OX::h[n - 1 - i];
  if (j < OY::M) hy = OY::h[j];
  else if (j >= m - OY::M) hy = OY::h[m - 1 - j];
  const double fx = OX::is_h2 ? 1.0 / (double)(n - 2) : 1.0 / (double)(n - 1);
  const double fy = 1.0 / (OY::is_h2 ? (double)(m - 2) : (double)(m - 1));
  return (hx * fx) * (hy * fy);
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
// The density and pressure checks of upwind_dissipation (lib.rs:636-646) run
// when `check` is set: the stage kernel sets it at the points a tile owns in
// its ξ phase, which cover every point of a block once.
template <bool UP>
__device__ __forceinline__ void face_terms(double rho, double ru, double rv, double e, double jx, double jy, double gamma,
                                           double f[4], double aq[4], double& alpha, unsigned& err, bool check = true) {
  const double ir = frcp(rho);
  const double u = ru * ir, v = rv * ir;
  const double p = (gamma - 1.0) * (e - (ru * ru + rv * rv) * (0.5 * ir));
  const double uh = jx * u + jy * v;
  f[0] = rho * uh;
  f[1] = ru * uh + jx * p;
  f[2] = rv * uh + jy * p;
  f[3] = (e + p) * uh;
  if constexpr (UP) {
    if (check) {
      if (!(rho > 0.0)) err |= ERRBIT_RHO;
      if (!(p > 0.0)) err |= ERRBIT_PRESSURE;
    }
    const double c2g = gamma * p * ir * (jx * jx + jy * jy);
    const double al = fabs(uh) + fsqrt(fmax(c2g, 1e-300));
    alpha = al;
    aq[0] = al * rho;
    aq[1] = al * ru;
    aq[2] = al * rv;
    aq[3] = al * e;
  }
}

// ---------------------------------------------------------------------------
// Mixed form (SBP_MIXED): the stencils run in FP32 on deviations of the face
// terms from those of a reference state q_C that is constant over one tile
// (its centre point), so FP32 rounding acts on local variations, not on the
// magnitudes of the state. The stencils of both operators have zero row
// sums, so D(α q) = D(α q - α_C q_C) exactly, and for the fluxes
// Ê = J·ξx E + J·ξy F
//   D_ξ Ê + D_η F̂ = D_ξ(Ê - Ê_C) + D_η(F̂ - F̂_C)
//                   + E_C (D_ξ J·ξx + D_η J·ηx) + F_C (D_ξ J·ξy + D_η J·ηy),
// with Ê_C = J·ξx E(q_C) + J·ξy F(q_C) at each point's metrics; the last two
// terms vanish identically for metrics computed with the same operators
// (J·ξx = D_η y, J·ηx = -D_ξ y, and D_ξ, D_η commute), so the mixed kernel
// differentiates only the deviations.
// The flux and dissipation terms of a point are evaluated in FP64
// (face_terms) and only their deviations from the reference terms are
// rounded to FP32:
//   f  = Ê(q; jx, jy) - (jx E(q_C) + jy F(q_C))   (Ê_C with the point's metrics)
//   aq = α q - α_C q_C                              (α_C with the centre metrics)
// The subtraction of two nearby FP64 values leaves the deviation with an
// absolute error of a few ulps of the terms; its rounding to FP32 then acts
// on the deviation only. (An earlier form derived the deviations with FP32
// algebra from δq = q - q_C: 12 % more instructions, no faster, and a
// 1.15x larger 1000-step error against the FP64 form.)
struct MixRef64 {
  double ec[4], fc[4];  // Cartesian fluxes E(q_C), F(q_C)
  double aqc[4];        // α_C q_C (α_C with the phase's metric pair at the centre)
};

// Reference terms of one tile in the per-tile layout of StageArgs::mixref
// (16 doubles): E(q_C) at [0, 4), F(q_C) at [4, 8), α_ξ,C q_C at [8, 12)
// (α with the centre's J·∇ξ), α_η,C q_C at [12, 16) (J·∇η). The mixed form
// evaluates them once per tile and RK4 step, from y_n at the tile's centre
// (mix_ref_kernel): any state the whole tile shares drops out of the
// stencils, so only its closeness to the tile's stage inputs matters, and
// the stage input moves by O(dt) within a step.
__device__ __forceinline__ void mix_ref_terms(const double q[4], double jxx, double jxy, double jex, double jey,
                                              double gamma, double r[16]) {
  const double ir = frcp(q[0]);
  const double u = q[1] * ir, v = q[2] * ir;
  const double p = (gamma - 1.0) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  r[0] = q[1];
  r[1] = q[1] * u + p;
  r[2] = q[2] * u;
  r[3] = (q[3] + p) * u;
  r[4] = q[2];
  r[5] = q[1] * v;
  r[6] = q[2] * v + p;
  r[7] = (q[3] + p) * v;
  const double c2 = gamma * p * ir;
  const double ax = fabs(jxx * u + jxy * v) + fsqrt(fmax(c2 * (jxx * jxx + jxy * jxy), 1e-300));
  const double ae = fabs(jex * u + jey * v) + fsqrt(fmax(c2 * (jex * jex + jey * jey), 1e-300));
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    r[8 + c] = ax * q[c];
    r[12 + c] = ae * q[c];
  }
}

// The reference terms of one phase (0: ξ, 1: η) from a tile's 16 doubles;
// the address is the same for the whole workgroup, so these are scalar loads.
__device__ __forceinline__ MixRef64 load_mix_ref(const double* __restrict__ rf, int phase) {
  MixRef64 R;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    R.ec[c] = rf[c];
    R.fc[c] = rf[4 + c];
    R.aqc[c] = rf[8 + 4 * phase + c];
  }
  return R;
}

template <bool UP>
__device__ __forceinline__ void face_terms_mix64(const double q[4], double jx, double jy, const MixRef64& R, double gamma,
                                                 float f[4], float aq[4], unsigned& err, bool check = true) {
  double f64[4], aq64[4], al;
  face_terms<UP>(q[0], q[1], q[2], q[3], jx, jy, gamma, f64, aq64, al, err, check);
#pragma unroll
  for (int c = 0; c < 4; ++c) f[c] = (float)(f64[c] - (jx * R.ec[c] + jy * R.fc[c]));
  if constexpr (UP) {
#pragma unroll
    for (int c = 0; c < 4; ++c) aq[c] = (float)(aq64[c] - R.aqc[c]);
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

// FP64 stage inputs: y_n, Y2 (BUF_YA), Y3 (BUF_YB), Y4 (BUF_YA again: Y2 is
// dead once stage 2 has formed W into BUF_ACC); the right-hand-side mode
// reads rhs_in or the selected work set.
// Y2, Y3, Y4 and W are internal to a step and stored point-major:
// the four components of point o at doubles 4o .. 4o + 3, two 16-byte loads
// per point; y_n and every buffer a caller sees keep the plane layout.
__device__ __forceinline__ bool stage_input_pm(const StageArgs& A, int mode) {
  return !A.mixed && ((mode >= MODE_S1 && mode <= MODE_S3) || mode == MODE_BS1 || mode == MODE_BS2);
}
__device__ __forceinline__ void load_pm(const double* base, uint32_t o, double q[4]) {
  const d2* p = reinterpret_cast<const d2*>(base);
  const d2 a = ldg(p, 2 * o), b = ldg(p, 2 * o + 1);
  q[0] = a.x;
  q[1] = a.y;
  q[2] = b.x;
  q[3] = b.y;
}
__device__ __forceinline__ void load_pm_nt(const double* base, uint32_t o, double q[4]) {
  const d2* p = reinterpret_cast<const d2*>(base);
  const d2 a = ldg_nt(p, 2 * o), b = ldg_nt(p, 2 * o + 1);
  q[0] = a.x;
  q[1] = a.y;
  q[2] = b.x;
  q[3] = b.y;
}
__device__ __forceinline__ void store_pm_nt(double* base, uint32_t o, double q0, double q1, double q2, double q3) {
  d2* p = reinterpret_cast<d2*>(base);
  d2 a, b;
  a.x = q0;
  a.y = q1;
  b.x = q2;
  b.y = q3;
  stg_nt(p, 2 * o, a);
  stg_nt(p, 2 * o + 1, b);
}
__device__ __forceinline__ const double* stage_input(const StageArgs& A, const BlockDev& B, int mode) {
  switch (mode) {
    case MODE_S0: return B.buf[BUF_YN];
    case MODE_S1: return B.buf[BUF_YA];
    case MODE_S2: return B.buf[BUF_YB];
    case MODE_S3: return B.buf[BUF_YA];
    // embedded Bogacki-Shampine stages: y_n, Y2, Y3, the solution
    case MODE_BS0: return A.swap ? B.buf[BUF_ACC] : B.buf[BUF_YN];
    case MODE_BS1: return B.buf[BUF_YA];
    case MODE_BS2: return B.buf[BUF_YB];
    case MODE_BS3: return A.swap ? B.buf[BUF_YN] : B.buf[BUF_ACC];
    default: return A.rhs_in_sel ? B.rk[A.rhs_in_sel - 1] : B.rhs_in;
  }
}

// Buffers of the mixed form (see BufRole).
__device__ __forceinline__ double* yn_buf(const BlockDev& B, const StageArgs& A) {
  return A.swap ? B.buf[BUF_ACC] : B.buf[BUF_YN];
}
__device__ __forceinline__ double* acc_buf(const BlockDev& B, const StageArgs& A) {
  return A.swap ? B.buf[BUF_YN] : B.buf[BUF_ACC];
}
// Stage-increment sets of the mixed form: Δ_1 (set 0) and Δ_2 (set 1) in the
// YA buffer, Δ_3 (set 2) in the first half of YB. Stage s (S0..S2) writes
// set s; stage s (S1..S3) reads set s - 1 for its input y_n + Δ_s. A set
// holds the four components of a point together (one 16-byte load or store
// per point), at the point's element offset o of the plane layout.

__device__ __forceinline__ const f4* delta_set(const BlockDev& B, int set) {
  return reinterpret_cast<const f4*>(reinterpret_cast<const float*>(B.buf[set == 2 ? BUF_YB : BUF_YA]) +
                                     (set == 1 ? 4 * B.ps : 0));
}
__device__ __forceinline__ const f4* delta_in(const BlockDev& B, int mode) { return delta_set(B, mode - 1); }
__device__ __forceinline__ f4* delta_out(const BlockDev& B, int mode) { return const_cast<f4*>(delta_set(B, mode)); }

// The state stage `mode` reads at element offset o.
__device__ __forceinline__ void load_stage_state(const StageArgs& A, const BlockDev& B, int mode, uint32_t o, double q[4]) {
  const uint32_t ps = (uint32_t)B.ps;
  if (A.mixed && mode >= MODE_S0 && mode <= MODE_S3) {
    const double* yn = yn_buf(B, A);
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = ldg(yn + c * ps, o);
    if (mode >= MODE_S1) {
      const f4 d = ldg(delta_in(B, mode), o);
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] += (double)d[c];
    }
  } else if (stage_input_pm(A, mode)) {
    load_pm(stage_input(A, B, mode), o, q);
  } else {
    const double* in = stage_input(A, B, mode);
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = ldg(in + c * ps, o);
  }
}

// Time of block B at the current stage (the system time plus c_s·dt, scaled
// by the block's step scale).
__device__ __forceinline__ double stage_time(const StageArgs& A, const BlockDev& B) {
  return ((A.time ? *A.time : 0.0) + A.time_offset) * B.dt_scale;
}

// Boundary data of an Eval edge (boundary_extract's Eval, lib.rs:853-866):
// the expressions at the stage time on edge point o's coordinates.
__device__ __forceinline__ void eval_edge_point(const StageArgs& A, const BlockDev& B, int dir, uint32_t o, double g[4],
                                                unsigned& err) {
  if (!expr::eval_point(B.eprog[dir], stage_time(A, B), B.xc[o], B.yc[o], g)) {
    err |= ERRBIT_EXPR;
    g[0] = g[1] = g[2] = g[3] = 0.0;
  }
}

// Boundary data g of edge `dir` at edge index t (boundary_extracts, lib.rs:869-915).
// Offset of point m of the edge of block S opposite to `dir` (the source
// edge a neighbour reads: boundary_extracts' seldir, lib.rs:869-915).
__device__ __forceinline__ uint32_t source_edge_offset(const BlockDev& S, int dir, int m) {
  switch (dir) {
    case SBP_NORTH: return (uint32_t)m;                                   // source south row
    case SBP_SOUTH: return (uint32_t)((S.ny - 1) * S.ld + m);            // source north row
    case SBP_EAST: return (uint32_t)(m * S.ld);                           // source west column
    default: return (uint32_t)(m * S.ld + (S.nx - 1));                    // source east column
  }
}

// `o` is the element offset of edge point t in the block (for the
// coordinates of far-field edges); errors of the far-field data (a negative
// vortex density or pressure, an expression that fails) go to `err`.
__device__ __forceinline__ void edge_data(const StageArgs& A, const BlockDev& B, int dir, int t, uint32_t o, int mode,
                                          double g[4], unsigned& err) {
  const int src = B.edge_src[dir];
  if (src == EDGE_THIS || src == EDGE_GRID) {
    const BlockDev& S = src == EDGE_THIS ? B : A.blocks[B.edge_nb[dir]];
    load_stage_state(A, S, mode, source_edge_offset(S, dir, t), g);
  } else if (src == EDGE_VORTEX) {
    // the vortex at this stage's time, t·dt_scale for a block stepping with
    // its own dt (boundary_extract's Vortex, lib.rs:851-852, vortexify :983-998)
    g[0] = g[1] = g[2] = g[3] = 0.0;
    vortex_point(B.vortex[dir], A.gamma, stage_time(A, B), B.xc[o], B.yc[o], g, err);
  } else if (src == EDGE_EVAL) {
    // evaluated for this stage by the eval-edges kernel (eval_edge_point)
    const int len = (dir == SBP_NORTH || dir == SBP_SOUTH) ? B.nx : B.ny;
#pragma unroll
    for (int c = 0; c < 4; ++c) g[c] = ldg(B.edata[dir], (uint32_t)(c * len + t));
  } else if (src == EDGE_INTERP) {
    // Interpolate (lib.rs:816-828): the neighbour's opposite edge through a
    // 2:1 operator (the host checked the lengths)
    const BlockDev& S = A.blocks[B.edge_nb[dir]];
    const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
    const interp::Table T = interp::table(B.edge_iop[dir], B.edge_f2c[dir] != 0);
    interp::value<4>(T, t, ns ? B.nx : B.ny, ns ? S.nx : S.ny,
                     [&](int m, double v[4]) { load_stage_state(A, S, mode, source_edge_offset(S, dir, m), v); }, g);
  } else if (src == EDGE_MULTIGRID) {
    // MultiGrid (lib.rs:829-849): piece k covers [i_k, i_k + n_k) with
    // i_{k+1} = i_k + n_k - 1; a point takes the last piece covering it,
    // halved when it is that piece's first point and the piece is not the
    // last (the host checked that the pieces cover the edge)
    const int* pc = B.edge_pieces[dir];
    const int np = B.edge_npieces[dir];
    int i = 0, sel = 0, sel_i = 0;
    for (int k = 0; k < np; ++k) {
      const int len = pc[3 * k + 2] - pc[3 * k + 1];
      if (t >= i && t < i + len) {
        sel = k;
        sel_i = i;
      }
      i += len - 1;
    }
    const BlockDev& S = A.blocks[pc[3 * sel]];
    load_stage_state(A, S, mode, source_edge_offset(S, dir, pc[3 * sel + 1] + (t - sel_i)), g);
    if (t == sel_i && sel != np - 1) {
#pragma unroll
      for (int c = 0; c < 4; ++c) g[c] /= 2.0;
    }
  } else {
    const double* p = B.edge_ptr[dir];
#pragma unroll
    for (int c = 0; c < 4; ++c) g[c] = p[c * B.edge_cs[dir] + t];
  }
}

}  // namespace sbp
