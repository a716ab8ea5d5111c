// The fused Euler RK4 stage kernel (FP64).
//
// One workgroup computes the right-hand side of one TX x TY tile of one block
// and folds it straight into the RK4 stage combination, replacing in a single
// pass what the reference does with ~137 plane passes per stage:
//   fluxes + metric transform          euler/src/lib.rs:680-735
//   diffxi / diffeta (DIFF)            sbp/src/operators/algos.rs:101-187, :371-519
//   upwind dissipation + dissxi/eta    euler/src/lib.rs:619-674
//   k = (-dE - dF + ad_xi + ad_eta)/J  euler/src/lib.rs:606-613 (trad: :507-512)
//   SAT_characteristics                euler/src/lib.rs:1002-1252
//   RK4 stage combination              utils/integrate/src/lib.rs:188-221
//
// Schedule inside a workgroup (LDS reused by both phases):
//   ξ-phase: every point of the tile rows widened by the ξ halo computes its
//     face terms with (J·ξx, J·ξy); the flux goes to LDS, the dissipation
//     weight·state stays in registers; ξ-stencils of the flux (DIFF), then of
//     the weighted state (DISS) are read back from LDS.
//   η-phase: the same over the tile columns widened by the η halo, with
//     (J·ηx, J·ηy), and η-stencils.
//   finish: k = (idx_ξ·Sξ + idx_η·Sη)/J, edge SATs, RK4 combination, store.
// Threads own one ξ-column and RY consecutive η-points of the tile.
#pragma once
#include "sbp_common.hpp"

namespace sbp {

template <class OX, class OY, int TX, int TY, int RY>
struct StageGeom {
  static constexpr int NT = TX * (TY / RY);
  static constexpr int HX = OX::H, HY = OY::H;
  static constexpr int WX = TX + 2 * HX;  // ξ-phase region width
  static constexpr int HYR = TY + 2 * HY;  // η-phase region height
  static constexpr int NPX = TY * WX;
  static constexpr int NPY = HYR * TX;
  static constexpr int NP = NPX > NPY ? NPX : NPY;
  static constexpr int ITX = (NPX + NT - 1) / NT;
  static constexpr int ITY = (NPY + NT - 1) / NT;
  static constexpr int LDS_BYTES = 4 * NP * 8;
};

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

template <class OX, class OY, bool UP, int TX, int TY, int RY>
__global__ __launch_bounds__(TX * (TY / RY)) void stage_kernel(const StageArgs A, const int mode) {
  using G = StageGeom<OX, OY, TX, TY, RY>;
  constexpr int NT = G::NT, HX = G::HX, HY = G::HY, WX = G::WX, NP = G::NP;
  __shared__ double lds[4 * NP];

  const TileDev T = A.tiles[blockIdx.x];
  const BlockDev& B = A.blocks[T.block];
  const int nx = B.nx, ny = B.ny, ld = B.ld;
  const int64_t ps = B.ps;
  int x0, wx0, wx1, y0, wy0, wy1;
  tile_place(T.tx, B.ntx, nx, TX, OX::M, x0, wx0, wx1);
  tile_place(T.ty, B.nty, ny, TY, OY::M, y0, wy0, wy1);
  const double* __restrict__ in = stage_input(B, mode);
  const double* __restrict__ met = B.metrics;
  const double gamma = A.gamma;
  unsigned err = 0;

  const int tid = threadIdx.x;
  const int tcol = tid % TX, trow = tid / TX;
  const int x = x0 + tcol;
  const int yb = y0 + trow * RY;
  const int yend = min(y0 + TY, ny);
  // Only the points this tile writes are computed: every stencil they need
  // lies inside the loaded regions (see tile_place).
  const bool xin = x >= wx0 && x < wx1;

  double acc[RY][4];

  // ------------------------------------------------------------ ξ-phase
  {
    using BD = Blk<OX, SBP_DIFF>;
    const int rx0 = max(0, x0 - HX), rx1 = min(nx, x0 + TX + HX);
    const int xw = rx1 - rx0;
    const bool interior = x0 >= OX::M && x0 + TX <= nx - OX::M;
    double aq[G::ITX][4];
#pragma unroll
    for (int it = 0; it < G::ITX; ++it) {
      const int l = tid + it * NT;
      const int r = l / WX, c = l - r * WX;
      const int gy = y0 + r;
      if (l < G::NPX && c < xw && gy < yend) {
        const int64_t o = (int64_t)gy * ld + (rx0 + c);
        double f[4];
        face_terms<UP>(in[o], in[ps + o], in[2 * ps + o], in[3 * ps + o], met[ps + o], met[2 * ps + o], gamma, f, aq[it],
                       err);
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < RY; ++s) {
      const int yl = trow * RY + s;
      if (xin && yb + s >= wy0 && yb + s < wy1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double* line = lds + q * NP + yl * WX;
          acc[s][q] = -(interior ? interior_sum<BD>(line + (x - rx0), 1) : row_sum<BD>(line, 1, x, nx, rx0));
        }
      }
    }
    if constexpr (UP) {
      using BS = Blk<OX, SBP_DISS>;
      __syncthreads();
#pragma unroll
      for (int it = 0; it < G::ITX; ++it) {
        const int l = tid + it * NT;
        const int r = l / WX, c = l - r * WX;
        if (l < G::NPX && c < xw && y0 + r < yend) {
#pragma unroll
          for (int q = 0; q < 4; ++q) lds[q * NP + l] = aq[it][q];
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < RY; ++s) {
        const int yl = trow * RY + s;
        if (xin && yb + s >= wy0 && yb + s < wy1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double* line = lds + q * NP + yl * WX;
            acc[s][q] += interior ? interior_sum<BS>(line + (x - rx0), 1) : row_sum<BS>(line, 1, x, nx, rx0);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < RY; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[s][q] *= B.idx_xi;
    __syncthreads();
  }

  // ------------------------------------------------------------ η-phase
  {
    using BD = Blk<OY, SBP_DIFF>;
    const int ry0 = max(0, y0 - HY), ry1 = min(ny, y0 + TY + HY);
    const int yh = ry1 - ry0;
    const int xend = min(x0 + TX, nx);
    const bool interior = y0 >= OY::M && y0 + TY <= ny - OY::M;
    double aq[G::ITY][4];
#pragma unroll
    for (int it = 0; it < G::ITY; ++it) {
      const int l = tid + it * NT;
      const int r = l / TX, c = l - r * TX;
      if (l < G::NPY && r < yh && x0 + c < xend) {
        const int64_t o = (int64_t)(ry0 + r) * ld + (x0 + c);
        double f[4];
        face_terms<UP>(in[o], in[ps + o], in[2 * ps + o], in[3 * ps + o], met[3 * ps + o], met[4 * ps + o], gamma, f, aq[it],
                       err);
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
      }
    }
    __syncthreads();
    double sy[RY][4];
#pragma unroll
    for (int s = 0; s < RY; ++s) {
      const int y = yb + s;
      if (xin && y >= wy0 && y < wy1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double* col = lds + q * NP + tcol;
          sy[s][q] = -(interior ? interior_sum<BD>(col + (y - ry0) * TX, TX) : row_sum<BD>(col, TX, y, ny, ry0));
        }
      }
    }
    if constexpr (UP) {
      using BS = Blk<OY, SBP_DISS>;
      __syncthreads();
#pragma unroll
      for (int it = 0; it < G::ITY; ++it) {
        const int l = tid + it * NT;
        const int r = l / TX, c = l - r * TX;
        if (l < G::NPY && r < yh && x0 + c < xend) {
#pragma unroll
          for (int q = 0; q < 4; ++q) lds[q * NP + l] = aq[it][q];
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < RY; ++s) {
        const int y = yb + s;
        if (xin && y >= wy0 && y < wy1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double* col = lds + q * NP + tcol;
            sy[s][q] += interior ? interior_sum<BS>(col + (y - ry0) * TX, TX) : row_sum<BS>(col, TX, y, ny, ry0);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < RY; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[s][q] += B.idx_eta * sy[s][q];
  }

  // ------------------------------------------------------------ finish
  const double time = (A.time ? *A.time : 0.0) + A.time_offset;
#pragma unroll
  for (int s = 0; s < RY; ++s) {
    const int y = yb + s;
    if (!(xin && y >= wy0 && y < wy1)) continue;
    const int64_t o = (int64_t)y * ld + x;
    const double ij = 1.0 / met[o];
    double k[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) k[q] = acc[s][q] * ij;
    if (x == 0 || x == nx - 1 || y == 0 || y == ny - 1) {
      double qv[4], g[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) qv[q] = in[q * ps + o];
      if (y == ny - 1) {
        edge_data(A, B, SBP_NORTH, x, x, y, mode, time, g, err);
        sat_point(k, qv, g, B.hi_eta, -1.0, 1.0, met[o], met[3 * ps + o], met[4 * ps + o], gamma);
      }
      if (y == 0) {
        edge_data(A, B, SBP_SOUTH, x, x, y, mode, time, g, err);
        sat_point(k, qv, g, B.hi_eta, 1.0, -1.0, met[o], met[3 * ps + o], met[4 * ps + o], gamma);
      }
      if (x == nx - 1) {
        edge_data(A, B, SBP_EAST, y, x, y, mode, time, g, err);
        sat_point(k, qv, g, B.hi_xi, -1.0, 1.0, met[o], met[ps + o], met[2 * ps + o], gamma);
      }
      if (x == 0) {
        edge_data(A, B, SBP_WEST, y, x, y, mode, time, g, err);
        sat_point(k, qv, g, B.hi_xi, 1.0, -1.0, met[o], met[ps + o], met[2 * ps + o], gamma);
      }
    }
    if (mode == MODE_RHS) {
#pragma unroll
      for (int q = 0; q < 4; ++q) B.rhs_out[q * ps + o] = k[q];
    } else if (mode == MODE_S0) {
      double* accb = B.buf[BUF_ACC];
      double* nxt = B.buf[BUF_YA];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double yn = in[q * ps + o];
        accb[q * ps + o] = yn + A.rk_scale_acc * k[q];
        nxt[q * ps + o] = yn + A.rk_scale_next * k[q];
      }
    } else if (mode == MODE_S3) {
      const double* accb = B.buf[BUF_ACC];
      double* out = B.buf[BUF_YN];
#pragma unroll
      for (int q = 0; q < 4; ++q) out[q * ps + o] = accb[q * ps + o] + A.rk_scale_acc * k[q];
    } else {
      const double* ynb = B.buf[BUF_YN];
      double* accb = B.buf[BUF_ACC];
      double* nxt = B.buf[mode == MODE_S1 ? BUF_YB : BUF_YA];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double yn = ynb[q * ps + o];
        accb[q * ps + o] = accb[q * ps + o] + A.rk_scale_acc * k[q];
        nxt[q * ps + o] = yn + A.rk_scale_next * k[q];
      }
    }
  }
  if (err) atomicOr(A.error, err);
}

}  // namespace sbp
