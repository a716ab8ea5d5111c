// The fused Euler RK4 stage kernel (FP64).
//
// One workgroup computes the right-hand side of one TX x TY tile of one block
// and folds it straight into the RK4 stage combination, replacing in a single
// pass what the reference does with ~137 plane passes per stage:
//   fluxes + metric transform          euler/src/lib.rs:680-735
//   diffxi / diffeta (DIFF)            sbp/src/operators/algos.rs:101-187, :371-519
//   upwind dissipation + dissxi/eta    euler/src/lib.rs:619-674
//   k = (-dE - dF + ad_xi + ad_eta)/J  euler/src/lib.rs:606-613 (trad: :507-512)
//   RK4 stage combination              utils/integrate/src/lib.rs:188-221
//
// Schedule inside a workgroup (LDS reused by both phases):
//   ξ-phase: every point of the tile rows widened by the ξ halo computes its
//     face terms with (J·ξx, J·ξy); the flux goes to LDS, the dissipation
//     weight·state stays in registers; ξ-stencils of the flux (DIFF), then of
//     the weighted state (DISS) are read back from LDS.
//   η-phase: the same over the tile columns widened by the η halo, with
//     (J·ηx, J·ηy), and η-stencils.
//   finish: k = (idx_ξ·Sξ + idx_η·Sη)/J, RK4 combination, store.
// The characteristic SATs of the edge points run afterwards in
// sat_edge_kernel: the stage combination is linear in k, so the SAT term
// enters as acc += b·dt·sat and next += a·dt·sat. Keeping the 4x4
// characteristic algebra out of this kernel keeps its register count low.
// Threads own one ξ-column and RY consecutive η-points of the tile.
#pragma once
#include "sbp_common.hpp"

namespace sbp {

template <class OX, class OY, int TX, int TY, int RY, int VAR = 0>
struct StageGeom {
  static constexpr int NT = TX * (TY / RY);
  // values per point held in LDS at once: 4 (two passes) or 8 (one pass)
  static constexpr int NV = VAR == 2 ? 8 : 4;
  static constexpr int HX = OX::H, HY = OY::H;
  static constexpr int WX = TX + 2 * HX;  // ξ-phase region width
  static constexpr int HYR = TY + 2 * HY;  // η-phase region height
  static constexpr int NPX = TY * WX;
  static constexpr int NPY = HYR * TX;
  static constexpr int NP = NPX > NPY ? NPX : NPY;
  static constexpr int ITX = (NPX + NT - 1) / NT;
  static constexpr int ITY = (NPY + NT - 1) / NT;
  static constexpr int LDS_BYTES = NV * NP * 8;
};

// Variants (autotuned): VAR 0 keeps the dissipation terms α·q of a thread's
// region points in registers between the flux and dissipation passes; VAR 1
// keeps only α and re-reads q (cache hits); VAR 2 stores flux and α·q
// together (8 values per point in LDS, one pass per direction).
template <class OX, class OY, bool UP, int TX, int TY, int RY, int VAR>
__global__ __launch_bounds__(TX * (TY / RY)) void stage_kernel(const StageArgs A, const int mode) {
  using G = StageGeom<OX, OY, TX, TY, RY, VAR>;
  constexpr int NT = G::NT, HX = G::HX, HY = G::HY, WX = G::WX, NP = G::NP;
  constexpr int NV = UP ? G::NV : 4;
  constexpr int NC = VAR == 0 ? 4 : 1;  // cached values per region point between passes
  __shared__ double lds[NV * NP];

  const TileDev T = A.tiles[blockIdx.x];
  const BlockDev& B = A.blocks[T.block];
  const int nx = B.nx, ny = B.ny, ld = B.ld;
  // 32-bit element offsets: the host keeps 5 planes of a block below 2^29 elements
  const uint32_t ps = (uint32_t)B.ps;
  int x0, wx0, wx1, y0, wy0, wy1;
  tile_place(T.tx, B.ntx, nx, TX, OX::M, x0, wx0, wx1);
  tile_place(T.ty, B.nty, ny, TY, OY::M, y0, wy0, wy1);
  const double* __restrict__ in = stage_input(B, mode);
  const double* __restrict__ met = B.metrics;
  const double gamma = A.gamma;
  const double idx_xi = B.idx_xi, idx_eta = B.idx_eta;
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
    double cache[G::ITX][NC];
    // all loads of the phase first (invalid slots read a clamped in-bounds
    // point), so their latencies overlap; then the arithmetic
    double v[G::ITX][6];
#pragma unroll
    for (int it = 0; it < G::ITX; ++it) {
      const int l = tid + it * NT;
      const int r = l / WX, c = l - r * WX;
      const bool ok = l < G::NPX && c < xw && y0 + r < yend;
      const uint32_t o = ok ? (uint32_t)((y0 + r) * ld + rx0 + c) : (uint32_t)(y0 * ld + rx0);
      v[it][0] = ldg(in, o);
      v[it][1] = ldg(in, ps + o);
      v[it][2] = ldg(in, 2 * ps + o);
      v[it][3] = ldg(in, 3 * ps + o);
      v[it][4] = ldg(met, ps + o);
      v[it][5] = ldg(met, 2 * ps + o);
    }
#pragma unroll
    for (int it = 0; it < G::ITX; ++it) {
      const int l = tid + it * NT;
      const int r = l / WX, c = l - r * WX;
      const int gy = y0 + r;
      if (l < G::NPX && c < xw && gy < yend) {
        double f[4], aq[4], al;
        face_terms<UP>(v[it][0], v[it][1], v[it][2], v[it][3], v[it][4], v[it][5], gamma, f, aq, al, err);
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
        if constexpr (UP) {
          if constexpr (VAR == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) cache[it][q] = aq[q];
          } else if constexpr (VAR == 1) {
            cache[it][0] = al;
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) lds[(4 + q) * NP + l] = aq[q];
          }
        }
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
      constexpr int QB = VAR == 2 ? 4 : 0;  // LDS slot of α·q
      if constexpr (VAR != 2) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < G::ITX; ++it) {
          const int l = tid + it * NT;
          const int r = l / WX, c = l - r * WX;
          const int gy = y0 + r;
          if (l < G::NPX && c < xw && gy < yend) {
            if constexpr (VAR == 0) {
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][q];
            } else {
              const uint32_t o = (uint32_t)(gy * ld + rx0 + c);
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * ldg(in, q * ps + o);
            }
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int s = 0; s < RY; ++s) {
        const int yl = trow * RY + s;
        if (xin && yb + s >= wy0 && yb + s < wy1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double* line = lds + (QB + q) * NP + yl * WX;
            acc[s][q] += interior ? interior_sum<BS>(line + (x - rx0), 1) : row_sum<BS>(line, 1, x, nx, rx0);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < RY; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[s][q] *= idx_xi;
    __syncthreads();
  }

  // ------------------------------------------------------------ η-phase
  {
    using BD = Blk<OY, SBP_DIFF>;
    const int ry0 = max(0, y0 - HY), ry1 = min(ny, y0 + TY + HY);
    const int yh = ry1 - ry0;
    const int xend = min(x0 + TX, nx);
    const bool interior = y0 >= OY::M && y0 + TY <= ny - OY::M;
    double cache[G::ITY][NC];
    double v[G::ITY][6];
#pragma unroll
    for (int it = 0; it < G::ITY; ++it) {
      const int l = tid + it * NT;
      const int r = l / TX, c = l - r * TX;
      const bool ok = l < G::NPY && r < yh && x0 + c < xend;
      const uint32_t o = ok ? (uint32_t)((ry0 + r) * ld + x0 + c) : (uint32_t)(ry0 * ld + x0);
      v[it][0] = ldg(in, o);
      v[it][1] = ldg(in, ps + o);
      v[it][2] = ldg(in, 2 * ps + o);
      v[it][3] = ldg(in, 3 * ps + o);
      v[it][4] = ldg(met, 3 * ps + o);
      v[it][5] = ldg(met, 4 * ps + o);
    }
#pragma unroll
    for (int it = 0; it < G::ITY; ++it) {
      const int l = tid + it * NT;
      const int r = l / TX, c = l - r * TX;
      if (l < G::NPY && r < yh && x0 + c < xend) {
        double f[4], aq[4], al;
        face_terms<UP>(v[it][0], v[it][1], v[it][2], v[it][3], v[it][4], v[it][5], gamma, f, aq, al, err);
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
        if constexpr (UP) {
          if constexpr (VAR == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) cache[it][q] = aq[q];
          } else if constexpr (VAR == 1) {
            cache[it][0] = al;
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) lds[(4 + q) * NP + l] = aq[q];
          }
        }
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
      constexpr int QB = VAR == 2 ? 4 : 0;
      if constexpr (VAR != 2) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < G::ITY; ++it) {
          const int l = tid + it * NT;
          const int r = l / TX, c = l - r * TX;
          if (l < G::NPY && r < yh && x0 + c < xend) {
            if constexpr (VAR == 0) {
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][q];
            } else {
              const uint32_t o = (uint32_t)((ry0 + r) * ld + x0 + c);
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * ldg(in, q * ps + o);
            }
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int s = 0; s < RY; ++s) {
        const int y = yb + s;
        if (xin && y >= wy0 && y < wy1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double* col = lds + (QB + q) * NP + tcol;
            sy[s][q] += interior ? interior_sum<BS>(col + (y - ry0) * TX, TX) : row_sum<BS>(col, TX, y, ny, ry0);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < RY; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[s][q] += idx_eta * sy[s][q];
  }

  // ------------------------------------------------------------ finish
  // (the edge SATs follow in sat_edge_kernel, added through the same linear
  // stage combination)
#pragma unroll
  for (int s = 0; s < RY; ++s) {
    const int y = yb + s;
    if (!(xin && y >= wy0 && y < wy1)) continue;
    const uint32_t o = (uint32_t)(y * ld + x);
    const double ij = frcp(ldg(met, o));
    double k[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) k[q] = acc[s][q] * ij;
    if (mode == MODE_RHS) {
#pragma unroll
      for (int q = 0; q < 4; ++q) stg(B.rhs_out, q * ps + o, k[q]);
    } else if (mode == MODE_S0) {
      double* accb = B.buf[BUF_ACC];
      double* nxt = B.buf[BUF_YA];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double yn = ldg(in, q * ps + o);
        stg(accb, q * ps + o, yn + A.rk_scale_acc * k[q]);
        stg(nxt, q * ps + o, yn + A.rk_scale_next * k[q]);
      }
    } else if (mode == MODE_S3) {
      const double* accb = B.buf[BUF_ACC];
      double* out = B.buf[BUF_YN];
#pragma unroll
      for (int q = 0; q < 4; ++q) stg(out, q * ps + o, ldg(accb, q * ps + o) + A.rk_scale_acc * k[q]);
    } else {
      const double* ynb = B.buf[BUF_YN];
      double* accb = B.buf[BUF_ACC];
      double* nxt = B.buf[mode == MODE_S1 ? BUF_YB : BUF_YA];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double yn = ldg(ynb, q * ps + o);
        stg(accb, q * ps + o, ldg((const double*)accb, q * ps + o) + A.rk_scale_acc * k[q]);
        stg(nxt, q * ps + o, yn + A.rk_scale_next * k[q]);
      }
    }
  }
  if (err) record_error(A.error, err);
}

}  // namespace sbp
