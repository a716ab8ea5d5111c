// Marching form of the fused FP64 Euler RK4 stage kernel.
//
// One workgroup owns a strip of TX columns over a segment of SL rows of one
// block and walks down it in chunks of R rows; thread (tc, tr) owns column
// x0 + tc of every chunk's row tr. Each row of the stage input and the metrics
// is loaded from HBM once (plus the ξ halo of the strip):
//   - the ξ face terms of a chunk (its rows, widened by the ξ halo) go to the
//     LDS buffer X and are differentiated along ξ at once;
//   - the η face terms of the strip's own columns go to an LDS ring of three
//     chunks, so that the η stencils of chunk c-1 (which reach HY rows into
//     chunks c-2 and c, and up to N rows for the closure blocks at the block
//     ends) are applied after chunk c is loaded;
//   - the ξ sums of a chunk wait in registers for that η pass, which finishes
//     the chunk with the stage combination.
// The next chunk's loads are issued before the current chunk's arithmetic.
// Same operator rows, face terms, stage combination and write ranges as the
// tiled kernel (sbp_stage.hpp); the segment's first chunk is preceded by the
// HY rows above it, and the last segment of a block is placed like the last
// tile (right-aligned, owning every end-closure row).
#pragma once
#include "sbp_stage.hpp"

namespace sbp {

// FP64 stage combination of one point (see the tiled kernel's finish). q is
// the stage input at the point; S1, S2 fa = y_n; S3 fa = Y2, fb = Y3, fc = y_n.
__device__ __forceinline__ void rk_operands_fp64(const BlockDev& B, int mode, uint32_t o, uint32_t ps, double fa[4],
                                                 double fb[4], double fc[4]) {
  if (mode == MODE_RHS || mode == MODE_S0) return;
  const double* pa = B.buf[mode == MODE_S3 ? BUF_YA : BUF_YN];
#pragma unroll
  for (int q = 0; q < 4; ++q) fa[q] = ldg_nt(pa, q * ps + o);
  if (mode == MODE_S3) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      fb[q] = ldg_nt((const double*)B.buf[BUF_YB], q * ps + o);
      fc[q] = ldg_nt((const double*)B.buf[BUF_YN], q * ps + o);
    }
  }
}

__device__ __forceinline__ void rk_combine_fp64(const BlockDev& B, int mode, uint32_t o, uint32_t ps, const double k[4],
                                                const double q4[4], const double fa[4], const double fb[4],
                                                const double fc[4], double sacc, double snext) {
  if (mode == MODE_RHS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) stg_nt(B.rhs_out, q * ps + o, k[q]);
  } else if (mode == MODE_S0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) stg_nt(B.buf[BUF_YA], q * ps + o, q4[q] + snext * k[q]);
  } else if (mode == MODE_S1 || mode == MODE_S2) {
    double* nxt = B.buf[mode == MODE_S1 ? BUF_YB : BUF_ACC];
#pragma unroll
    for (int q = 0; q < 4; ++q) stg_nt(nxt, q * ps + o, fa[q] + snext * k[q]);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      stg_nt(B.buf[BUF_YN], q * ps + o, (fa[q] + 2.0 * fb[q] + q4[q] - fc[q]) * (1.0 / 3.0) + sacc * k[q]);
  }
}

template <class OX, class OY, bool UP, int TX, int R, int SL>
__global__ __launch_bounds__(TX* R, 2) void march_kernel(const StageArgs A, const int mode) {
  constexpr int NT = TX * R;
  constexpr int HX = OX::H, HY = OY::H;
  constexpr int WX = TX + 2 * HX;
  constexpr int RING = 3 * R;
  constexpr int NH = 2 * HX * R;  // ξ-halo points per chunk
  constexpr int NVY = UP ? 8 : 4;
  static_assert(NH <= NT, "at most one ξ-halo point per thread");
  static_assert(HY <= R && OY::N <= 2 * R && OY::N <= OY::M + R, "η windows within three chunks");
  static_assert(SL % R == 0 && SL >= OY::N, "segments of whole chunks");
  __shared__ __align__(16) double X[4 * R * WX];      // ξ face terms of the current chunk
  __shared__ __align__(16) double Y[NVY * RING * TX];  // η face terms, ring of three chunks

  const TileDev T = A.tiles[blockIdx.x];
  const BlockDev& B = A.blocks[T.block];
  const int nx = B.nx, ny = B.ny, ld = B.ld;
  const uint32_t ps = (uint32_t)B.ps;
  int x0, wx0, wx1, y0, wy0, wy1;
  tile_place(T.tx, B.ntx, nx, TX, OX::M, x0, wx0, wx1);
  tile_place(T.ty, B.nty, ny, SL, OY::M, y0, wy0, wy1);
  const double* __restrict__ in = stage_input(B, mode);
  const double* __restrict__ met = B.metrics;
  const double gamma = A.gamma;
  const double idx_xi = B.idx_xi, idx_eta = B.idx_eta;
  const double dtb = A.dt * B.dt_scale;
  const double sacc = A.rk_b * dtb, snext = A.rk_a * dtb;
  unsigned err = 0;

  const int tid = threadIdx.x;
  const int tc = tid % TX, tr = tid / TX;
  const int xo = x0 + tc;  // own column
  const bool xok = xo < nx;
  const bool xw = xo >= wx0 && xo < wx1;
  // this thread's ξ-halo point of each chunk (threads tid < NH)
  const int hr = tid / (2 * HX), hc = tid % (2 * HX);
  const int hx = hc < HX ? x0 - HX + hc : x0 + TX + (hc - HX);
  const bool hok = tid < NH && hx >= 0 && hx < nx;
  const int rx0 = x0 - HX;  // absolute column of X's first column
  const bool interior_x = x0 >= OX::M && x0 + TX <= nx - OX::M;

  const int yend = min(y0 + SL, ny);
  const int nchunks = (yend - y0 + R - 1) / R;  // chunks 1..nchunks hold rows y0..yend-1
  const int cb = y0 - R;                        // first row of chunk 0 (rows above the segment)
  const int rlo = max(0, y0 - HY), rhi = min(ny, yend + HY);  // rows whose face terms are needed

  // raw operands of one chunk: own point (q, J, J·ξx, J·ξy, J·ηx, J·ηy) and halo point (q, J·ξx, J·ξy)
  auto load_chunk = [&](int c, double p[9], double h[6]) {
    const int r = cb + c * R + tr;
    const bool ok = xok && r >= rlo && r < rhi;
    const uint32_t o = ok ? (uint32_t)(r * ld + xo) : (uint32_t)(y0 * ld + x0);
#pragma unroll
    for (int q = 0; q < 4; ++q) p[q] = ldg(in, q * ps + o);
#pragma unroll
    for (int m = 0; m < 5; ++m) p[4 + m] = ldg(met, m * ps + o);
    const int rh = cb + c * R + hr;
    const bool okh = hok && rh >= rlo && rh < rhi;
    const uint32_t oh = okh ? (uint32_t)(rh * ld + hx) : (uint32_t)(y0 * ld + x0);
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = ldg(in, q * ps + oh);
    h[4] = ldg(met, ps + oh);
    h[5] = ldg(met, 2 * ps + oh);
  };

  double cur[9], curh[6], nxt[9], nxth[6];
  double sxi[4] = {0.0, 0.0, 0.0, 0.0}, jp = 1.0;  // ξ sums and J of the pending point
  double fa[4], fb[4], fc[4], qp[4];                // stage-combination operands of the pending point
  load_chunk(0, cur, curh);

  for (int c = 0; c <= nchunks + 1; ++c) {
    if (c <= nchunks) load_chunk(c + 1, nxt, nxth);
    const int r = cb + c * R + tr;                    // own row of chunk c
    const bool rok = xok && r >= rlo && r < rhi;
    const int rh = cb + c * R + hr;
    const bool hrok = hok && rh >= rlo && rh < rhi;
    const int slot = (c % 3) * R + tr;
    double aqo[4], aqh[4];
    // ---- face terms of chunk c
    if (rok) {
      double f[4], al;
      face_terms<UP>(cur[0], cur[1], cur[2], cur[3], cur[5], cur[6], gamma, f, aqo, al, err);
#pragma unroll
      for (int q = 0; q < 4; ++q) X[(q * R + tr) * WX + HX + tc] = f[q];
      double g[4], ag[4];
      face_terms<UP>(cur[0], cur[1], cur[2], cur[3], cur[7], cur[8], gamma, g, ag, al, err);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Y[(q * RING + slot) * TX + tc] = g[q];
        if constexpr (UP) Y[((4 + q) * RING + slot) * TX + tc] = ag[q];
      }
    }
    if (hrok) {
      double f[4], al;
      face_terms<UP>(curh[0], curh[1], curh[2], curh[3], curh[4], curh[5], gamma, f, aqh, al, err);
#pragma unroll
      for (int q = 0; q < 4; ++q) X[(q * R + hr) * WX + (hx - rx0)] = f[q];
    }
    __syncthreads();
    // ---- ξ stencils of chunk c's own point (rows of the segment only)
    // only written points are finished: their stencils lie in the loaded rows and columns
    const bool fin_c = c >= 1 && c <= nchunks && xw && r >= wy0 && r < wy1;
    double sn[4] = {0.0, 0.0, 0.0, 0.0};
    auto xi_pass = [&](auto blk, double sg) {
      using BB = decltype(blk);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double* line = X + (q * R + tr) * WX;
        double o;
        if (interior_x) {
          double w[1];
          row_window<BB, 1>(line + tc, w);
          o = w[0];
        } else {
          o = row_sum<BB>(line, 1, xo, nx, rx0);
        }
        sn[q] = fma_t(sg, o, sn[q]);
      }
    };
    if (fin_c) xi_pass(Blk<OX, SBP_DIFF>{}, -idx_xi);
    if constexpr (UP) {
      __syncthreads();
      if (rok) {
#pragma unroll
        for (int q = 0; q < 4; ++q) X[(q * R + tr) * WX + HX + tc] = aqo[q];
      }
      if (hrok) {
#pragma unroll
        for (int q = 0; q < 4; ++q) X[(q * R + hr) * WX + (hx - rx0)] = aqh[q];
      }
      __syncthreads();
      if (fin_c) xi_pass(Blk<OX, SBP_DISS>{}, idx_xi);
    }
    // ---- η stencils and stage combination of chunk c-1
    if (c >= 2) {
      const int j = r - R;  // own row of chunk c-1
      if (xw && j >= wy0 && j < wy1) {
        // ring slot of absolute row a: (a - cb) mod RING
        auto eta_sum = [&](auto blk, int plane) {
          using BB = decltype(blk);
          const double* col = Y + plane * RING * TX + tc;
          double s = 0.0;
          if (j < BB::M) {
#pragma unroll 1
            for (int k = 0; k < BB::N; ++k) {
              int sl = (k - cb) % RING;
              s += BB::start(j, k) * col[sl * TX];
            }
          } else if (j >= ny - BB::M) {
            const int rr = j - (ny - BB::M);
#pragma unroll 1
            for (int k = 0; k < BB::N; ++k) {
              const int sl = (ny - BB::N + k - cb) % RING;
              s += BB::end(rr, k) * col[sl * TX];
            }
          } else {
            int sl = (j - BB::H - cb) % RING;
#pragma unroll
            for (int d = 0; d < BB::D; ++d) {
              if (BB::diag(d) != 0.0) s += BB::diag(d) * col[sl * TX];
              sl = sl + 1 == RING ? 0 : sl + 1;
            }
          }
          return s;
        };
        double k[4];
        const double ij = frcp(jp);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double a = fma_t(-idx_eta, eta_sum(Blk<OY, SBP_DIFF>{}, q), sxi[q]);
          if constexpr (UP) a = fma_t(idx_eta, eta_sum(Blk<OY, SBP_DISS>{}, 4 + q), a);
          k[q] = a * ij;
        }
        rk_combine_fp64(B, mode, (uint32_t)(j * ld + xo), ps, k, qp, fa, fb, fc, sacc, snext);
      }
    }
    // chunk c's point becomes the pending one
#pragma unroll
    for (int q = 0; q < 4; ++q) sxi[q] = sn[q];
    jp = cur[4];
    if (fin_c) {
#pragma unroll
      for (int q = 0; q < 4; ++q) qp[q] = cur[q];
      rk_operands_fp64(B, mode, (uint32_t)(r * ld + xo), ps, fa, fb, fc);
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) cur[m] = nxt[m];
#pragma unroll
    for (int m = 0; m < 6; ++m) curh[m] = nxth[m];
    __syncthreads();
  }
  if (err) record_error(A.error, err);
}

}  // namespace sbp
