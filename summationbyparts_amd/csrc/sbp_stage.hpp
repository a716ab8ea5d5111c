// The fused Euler RK4 stage kernel (FP64, and the mixed FP32-RHS form).
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
// Mixed form (MIX): the face terms are FP64, rounded to FP32 as deviations
// from those of a per-tile reference state (face_terms_mix64; the reference
// terms come from mix_ref_kernel once per step), LDS and
// stencil sums are FP32, y_n and
// the RK4 sum stay FP64; each stage input is y_n + Δ with the FP32 increment
// Δ = a·dt·k stored by the previous stage, and the last stage sums y_n and
// the increments in FP64.
// The EDGE instantiation, which the host launches for the tiles on block
// edges (alone, or over all tiles of a group with the edge tiles first), also
// carries the closure rows and the characteristic SATs of the block-edge
// points: a prologue evaluates them into LDS before the phases begin (when
// few registers are live), the finish adds them to k.
#pragma once
#include <type_traits>

#include "sbp_common.hpp"

namespace sbp {

template <class OX, class OY, int TX, int TY, int RX, int RY, int VAR = 0, int VBYTES = 8>
struct StageGeom {
  static constexpr int NT = (TX / RX) * (TY / RY);
  static_assert(NT % 64 == 0, "whole wavefronts per workgroup");
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
  static constexpr int LDS_BYTES = NV * NP * VBYTES;
  static_assert(RX == 1 || RX == 2, "RX is 1 or 2");
  static_assert(TX % 2 == 0 && NP % 2 == 0, "16-byte LDS alignment of paired reads");
};

// Interior stencils of R consecutive outputs from one window of R + D - 1
// values: along ξ the window is contiguous (read in pairs when R == 2),
// along η it is a column of pairs (two adjacent ξ-columns).
template <class B, int R, typename T>
__device__ __forceinline__ void row_window(const T* p, T out[R]) {
  constexpr int W = R + B::D - 1;
  T w[W + 1];
  if constexpr (R == 2) {
    using T2 = std::conditional_t<sizeof(T) == 8, double2, float2>;
    const T2* p2 = reinterpret_cast<const T2*>(p);
#pragma unroll
    for (int k = 0; k < (W + 1) / 2; ++k) {
      const T2 v = p2[k];
      w[2 * k] = v.x;
      w[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] = p[k];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T s = T(0);
#pragma unroll
    for (int d = 0; d < B::D; ++d)
      if (B::diag(d) != 0.0) s += T(B::diag(d)) * w[r + d];
    out[r] = s;
  }
}

template <class B, int RS, int RX, typename T>
__device__ __forceinline__ void col_window(const T* p, int stride, T out[RS][RX]) {
  constexpr int W = RS + B::D - 1;
#pragma unroll
  for (int i = 0; i < RX; ++i)
#pragma unroll
    for (int r = 0; r < RS; ++r) out[r][i] = T(0);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    T v[RX];
    if constexpr (RX == 2) {
      using T2 = std::conditional_t<sizeof(T) == 8, double2, float2>;
      const T2 t = *reinterpret_cast<const T2*>(p + k * stride);
      v[0] = t.x;
      v[1] = t.y;
    } else {
      v[0] = p[k * stride];
    }
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int d = k - r;
      if (d >= 0 && d < B::D && B::diag(d) != 0.0) {
#pragma unroll
        for (int i = 0; i < RX; ++i) out[r][i] += T(B::diag(d)) * v[i];
      }
    }
  }
}

// Variants (autotuned): VAR 0 keeps the dissipation terms α·q of a thread's
// region points in registers between the flux and dissipation passes; VAR 1
// keeps only α and re-reads q (cache hits); VAR 2 stores flux and α·q
// together (8 values per point in LDS, one pass per direction). The mixed
// form uses VAR 2: its FP32 values fit both terms in the LDS the FP64 form
// needs for one, and it saves two barriers per phase and the register cache
// of α·q. (VAR 1 at 16x24 and 16x16 fits 5 waves per SIMD, at most
// 96 VGPRs, but its re-reads cost more than the occupancy gains: 6.2 and 5.8
// against 7.2 Gpoint-steps/s at 4096².)
// Threads own RX adjacent ξ-columns and RY consecutive η-rows of the tile.
// The edge instantiation asks for 4 waves per SIMD (without the hint the
// upwind9 edge tiles land 2 VGPRs above that budget, at 130).
// ADAPT: the instantiation of the embedded Bogacki-Shampine modes (BS0-BS3,
// FP64 only), whose finish holds two operand sets; the RK4 and
// right-hand-side modes run in the other one, which needs fewer registers
// (the FP64 interior kernel 100 -> 92 VGPRs: 5 instead of 4 waves per SIMD).
template <class OX, class OY, bool UP, int TX, int TY, int RX, int RY, int VAR, bool MIX = false, bool EDGE = false,
          bool ADAPT = false>
__global__ __launch_bounds__((TX / RX) * (TY / RY)) __attribute__((amdgpu_waves_per_eu(EDGE ? 4 : 1, 8))) void stage_kernel(
    const StageArgs A, const int mode) {
  static_assert(!MIX || VAR == 2, "the mixed form keeps flux and α·q together in LDS");
  static_assert(!(MIX && ADAPT), "the adaptive modes are FP64");
  using CT = std::conditional_t<MIX, float, double>;  // face terms, LDS, stencil sums
  using G = StageGeom<OX, OY, TX, TY, RX, RY, VAR, sizeof(CT)>;
  constexpr int NT = G::NT, HX = G::HX, HY = G::HY, WX = G::WX, NP = G::NP;
  constexpr int NV = UP ? G::NV : 4;
  constexpr int NC = VAR == 0 ? 4 : 1;  // cached values per region point between passes
  constexpr int LCH = 3;                // region points whose loads are in flight together
  // Both phases start with a thread's own output points, loaded once in the
  // ξ phase; their stage input serves the η phase and (FP64) the stage
  // combination without a reload
  constexpr bool OWNQ = RX == 1 && RY <= LCH;
  constexpr int ITYE = OWNQ ? RY + (2 * G::HY * TX + NT - 1) / NT : G::ITY;
  constexpr int ITXE = OWNQ ? RY + (2 * G::HX * TY + NT - 1) / NT : G::ITX;
  // load batches per phase: the ξ phase loads all region points of a thread
  // at once when they are at most four (one memory round trip), else LCH at
  // a time; the η phase at most three at once, else two at a time (its
  // batch of four held 48 VGPRs, the kernel's peak: 32x16 tiles went from
  // 100 to 94 VGPRs in FP64 and 100 to 84 in the mixed form, 5 waves per
  // SIMD instead of 4, +2.5 % and +4 %)
  constexpr int LBX = ITXE <= 4 ? ITXE : LCH;
  constexpr int LBY = ITYE <= 3 ? ITYE : 2;
  __shared__ __align__(16) CT lds[NV * NP];

  const TileDev T = A.tiles[blockIdx.x];
  const BlockDev& B = A.blocks[T.block];
  const int nx = B.nx, ny = B.ny, ld = B.ld;
  // 32-bit element offsets: the host keeps 5 planes of a block below 2^29 elements
  const uint32_t ps = (uint32_t)B.ps;
  int x0, wx0, wx1, y0, wy0, wy1;
  tile_place(T.tx, B.ntx, nx, TX, OX::M, x0, wx0, wx1);
  tile_place(T.ty, B.nty, ny, TY, OY::M, y0, wy0, wy1);
  // stage input: y_n (FP64 planes), a point-major FP64 stage input, or in the
  // mixed form y_n plus the FP32 increment set (4 floats per point)
  const bool mixs = MIX && mode >= MODE_S0 && mode <= MODE_S3;
  const double* __restrict__ in = mixs ? yn_buf(B, A) : stage_input(A, B, mode);
  const f4* __restrict__ din = (MIX && mode >= MODE_S1 && mode <= MODE_S3) ? delta_in(B, mode) : nullptr;
  const double* __restrict__ met = B.metrics;
  // metric pairs: one 16-byte load per region point and phase
  const d2* __restrict__ mxi = reinterpret_cast<const d2*>(B.mpair);
  const d2* __restrict__ meta = reinterpret_cast<const d2*>(B.mpair + 2 * B.ps);
  const double gamma = A.gamma;
  const double idx_xi = B.idx_xi, idx_eta = B.idx_eta;
  unsigned err = 0;

  const int tid = threadIdx.x;
  const int tc = tid % (TX / RX), tr = tid / (TX / RX);
  const int xb = x0 + tc * RX;  // first ξ-column of this thread
  const int yb = y0 + tr * RY;  // first η-row of this thread
  const int yend = min(y0 + TY, ny);
  // Only the points this tile writes are computed: every stencil they need
  // lies inside the loaded regions (see tile_place).
  bool xin[RX], yin[RY];
#pragma unroll
  for (int i = 0; i < RX; ++i) xin[i] = xb + i >= wx0 && xb + i < wx1;
#pragma unroll
  for (int s = 0; s < RY; ++s) yin[s] = yb + s >= wy0 && yb + s < wy1;
  bool any = false;
#pragma unroll
  for (int i = 0; i < RX; ++i)
#pragma unroll
    for (int s = 0; s < RY; ++s) any = any || (xin[i] && yin[s]);

  // Characteristic SATs of the block-edge points this tile writes
  // (SAT_characteristic, lib.rs:1150-1252, boundary data as boundary_extracts
  // gives it, lib.rs:804-915), evaluated before anything else is live (or,
  // with A.sat_buf, read from the edge-SAT kernel's buffers) and kept in LDS
  // for the finish: slots [0, TX) north, [TX, 2TX) south, [2TX, 2TX+TY) east,
  // [2TX+TY, 2TX+2TY) west, by the point's offset from the tile origin along
  // the edge; zero where an edge has no SAT.
  constexpr int NSAT = EDGE ? 2 * TX + 2 * TY : 1;
  __shared__ __align__(16) double satl[4 * NSAT];
  if constexpr (EDGE) {
    // the block's descriptor read through a lane broadcast (taken while every
    // lane is active): the values the boundary data and the SAT use then live
    // in vector registers, which are free here, and not in the scalar
    // registers the stencils use
    const BlockDev& BV = A.blocks[__shfl(T.block, 0)];
    if (wy1 == ny || wy0 == 0 || wx1 == nx || wx0 == 0) {
      for (int it = threadIdx.x; it < NSAT; it += NT) {
        int dir, t;
        bool ok;
        if (it < 2 * TX) {
          dir = it < TX ? SBP_NORTH : SBP_SOUTH;
          t = x0 + (it < TX ? it : it - TX);
          ok = (it < TX ? wy1 == ny : wy0 == 0) && t >= wx0 && t < wx1;
        } else {
          const int e = it - 2 * TX;
          dir = e < TY ? SBP_EAST : SBP_WEST;
          t = y0 + (e < TY ? e : e - TY);
          ok = (e < TY ? wx1 == nx : wx0 == 0) && t >= wy0 && t < wy1;
        }
        double sat[4] = {0.0, 0.0, 0.0, 0.0};
        if (ok && BV.edge_src[dir] != EDGE_NONE && A.sat_buf) {
          // evaluated for this stage by the edge-SAT kernel
          const int len = (dir == SBP_NORTH || dir == SBP_SOUTH) ? nx : ny;
#pragma unroll
          for (int c = 0; c < 4; ++c) sat[c] = ldg(BV.sat[dir], (uint32_t)(c * len + t));
        } else if (ok && BV.edge_src[dir] != EDGE_NONE) {
          const bool ns = dir == SBP_NORTH || dir == SBP_SOUTH;
          const int gx = ns ? t : (dir == SBP_EAST ? nx - 1 : 0);
          const int gy = ns ? (dir == SBP_NORTH ? ny - 1 : 0) : t;
          const uint32_t o = (uint32_t)(gy * ld + gx);
          double q[4], g[4];
          load_stage_state(A, BV, mode, o, q);
          edge_data(A, BV, dir, t, o, mode, g, err);
          // (σ, τ): north (-1, +1), south (+1, -1), east (-1, +1), west (+1, -1);
          // north/south use J·∇η, east/west J·∇ξ
          const double sign = (dir == SBP_NORTH || dir == SBP_EAST) ? -1.0 : 1.0;
          const int m0 = ns ? 3 : 1;
          const double* mv = BV.metrics;
          sat_point(sat, q, g, ns ? BV.hi_eta : BV.hi_xi, sign, -sign, ldg(mv, o), ldg(mv + m0 * ps, o),
                    ldg(mv + (m0 + 1) * ps, o), gamma);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) satl[4 * it + c] = sat[c];
      }
    }
  }

  // Closure coefficients (start blocks of DIFF and, for upwind, DISS; the ξ
  // operator's, then the η operator's unless it is the same) copied into LDS
  // by the tiles that write closure rows: their stencils read a coefficient
  // row per lane (row_sum4_tab).
  // Per operator: the start blocks (DIFF, then DISS), then the interior
  // diagonals padded by N - D zeros on each side (DIFF, then DISS), for the
  // wavefronts that hold closure and interior rows together.
  constexpr bool SAMEOP = std::is_same<OX, OY>::value;
  constexpr int NK = UP ? 2 : 1;
  constexpr int LPX = 2 * OX::N - OX::D, LPY = 2 * OY::N - OY::D;
  constexpr int TABX = NK * (OX::M * OX::N + LPX), TABY = NK * (OY::M * OY::N + LPY);
  constexpr int NTAB = EDGE ? (SAMEOP ? TABX : TABX + TABY) : 1;
  __shared__ double ctab[NTAB];
  const bool xclos = !(x0 >= OX::M && x0 + TX <= nx - OX::M);
  const bool yclos = !(y0 >= OY::M && y0 + TY <= ny - OY::M);
  if constexpr (EDGE) {
    if (xclos || yclos) {
      auto fill = [&](auto op, int base) {
        using OP = decltype(op);
        constexpr int MN = OP::M * OP::N, LP = 2 * OP::N - OP::D, PAD = OP::N - OP::D;
        for (int e = tid; e < NK * (MN + LP); e += NT) {
          double v;
          if (e < NK * MN) {
            const int kind = e / MN, r = (e % MN) / OP::N, k = e % OP::N;
            if constexpr (UP) {
              v = kind == 0 ? OP::diff_start[r][k] : OP::diss_start[r][k];
            } else {
              v = OP::diff_start[r][k];
            }
          } else {
            const int f = e - NK * MN, kind = f / LP, d = f % LP - PAD;
            if (d < 0 || d >= OP::D) {
              v = 0.0;
            } else if constexpr (UP) {
              v = kind == 0 ? OP::diff_diag[d] : OP::diss_diag[d];
            } else {
              v = OP::diff_diag[d];
            }
          }
          ctab[base + e] = v;
        }
      };
      fill(OX{}, 0);
      if constexpr (!SAMEOP) fill(OY{}, TABX);
    }
  }
  const double* tabx = ctab;
  const double* taby = ctab + (SAMEOP ? 0 : TABX);

  // state of the stage input at offset o (FP64; the mixed form adds Δ)
  const bool pm = !MIX && stage_input_pm(A, mode);  // point-major FP64 stage input
  // (dk: where to keep the mixed form's increment Δ of the point as well)
  auto load_q = [&](uint32_t o, double q[4], f4* dk = nullptr) {
    if (!MIX && pm) {
      load_pm(in, o, q);
      return;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = ldg(in + c * ps, o);
    if constexpr (MIX) {
      if (din) {
        const f4 d = ldg(din, o);
        if (dk) *dk = d;
#pragma unroll
        for (int c = 0; c < 4; ++c) q[c] += (double)d[c];
      }
    }
  };
  // mixed form: the tile's reference terms (mix_ref_kernel, this step's y_n
  // at the tile's centre)
  const double* __restrict__ mref = MIX ? A.mixref + 16 * (B.tile_off + (int64_t)T.ty * B.ntx + T.tx) : nullptr;

  CT acc[RY][RX][4];
  // metric products at the own output points (OWNQ), all four loaded in the
  // ξ phase: J·ηx, J·ηy stay in registers for the η phase, and
  // J = J·ξx J·ηy - J·ξy J·ηx (= y_η x_ξ - x_η y_ξ, sbp/src/grid.rs:110-131),
  // which replaces a load of J, is parked in LDS for the finish (each thread
  // reading back its own slots, to keep it out of registers).
  double ownm[OWNQ ? RY : 1][2];
  __shared__ double jsave[OWNQ ? RY * NT : 1];
  // stage input at the own output points (OWNQ): loaded in the ξ phase, used
  // again by the η phase (which loads only J·ηx, J·ηy there) and the finish
  double ownq[OWNQ ? RY : 1][4];
  // mixed form: the increment Δ_s of the stage input at the own points (the
  // last stage's sum needs Δ_3 apart from y_n + Δ_3)
  f4 ownd[MIX ? RY : 1];

  // ------------------------------------------------------------ ξ-phase
  {
    using BD = Blk<OX, SBP_DIFF>;
    const int rx0 = max(0, x0 - HX), rx1 = min(nx, x0 + TX + HX);
    const int xw = rx1 - rx0;
    const bool interior = x0 >= OX::M && x0 + TX <= nx - OX::M;
    MixRef64 R;
    if constexpr (MIX) R = load_mix_ref(mref, 0);
    // region point (tile row r, column c from rx0) of this thread's
    // iteration it: own output points first (OWNQ), then the halo columns
    auto xi_pt = [&](int it, int& r, int& c) -> bool {
      if constexpr (OWNQ) {
        if (it < RY) {
          r = tr * RY + it;
          c = xb - rx0;
          return y0 + r < yend && xb < nx;
        }
        const int h = tid + (it - RY) * NT;
        r = h / (2 * HX);
        const int hc = h - r * (2 * HX);
        const int ax = hc < HX ? x0 - HX + hc : x0 + TX + (hc - HX);
        c = ax - rx0;
        return h < 2 * HX * TY && y0 + r < yend && ax >= rx0 && ax < rx1;
      } else {
        const int l = tid + it * NT;
        r = l / WX;
        c = l - r * WX;
        return l < G::NPX && c < xw && y0 + r < yend;
      }
    };
    CT cache[ITXE][NC];
    // loads of up to LBX region points are issued together (invalid points
    // read a clamped in-bounds point; slots past ITXE load nothing) so that
    // their latencies overlap
#pragma unroll
    for (int i0 = 0; i0 < ITXE; i0 += LBX) {
      double v[LBX][6];
#pragma unroll
      for (int j = 0; j < LBX; ++j) {
        if (i0 + j >= ITXE) continue;
        int r, c;
        const bool ok = i0 + j < ITXE && xi_pt(i0 + j, r, c);
        const uint32_t o = ok ? (uint32_t)((y0 + r) * ld + rx0 + c) : (uint32_t)(y0 * ld + rx0);
        load_q(o, v[j], (MIX && i0 + j < RY) ? &ownd[MIX && i0 + j < RY ? i0 + j : 0] : nullptr);
        const d2 mx = ldg(mxi, o);
        v[j][4] = mx.x;
        v[j][5] = mx.y;
        // the own points' η metrics are issued with them (used by the η phase)
        if (OWNQ && i0 + j < RY) {
          const d2 me = ldg(meta, o);
          ownm[OWNQ ? i0 + j : 0][0] = me.x;
          ownm[OWNQ ? i0 + j : 0][1] = me.y;
        }
      }
      if constexpr (OWNQ) {
        if (i0 == 0) {
#pragma unroll
          for (int s2 = 0; s2 < RY; ++s2) {
            jsave[s2 * NT + tid] = v[s2][4] * ownm[s2][1] - v[s2][5] * ownm[s2][0];
#pragma unroll
            for (int q = 0; q < 4; ++q) ownq[s2][q] = v[s2][q];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < LBX; ++j) {
        const int it = i0 + j;
        int r, c;
        if (it < ITXE && xi_pt(it, r, c)) {
          const int l = r * WX + c;
          CT f[4], aq[4];
          double al = 0.0;
          // ρ and p are checked once per point: at the points this tile owns
          const bool chk = !OWNQ || it < RY;
          if constexpr (MIX) {
            face_terms_mix64<UP>(v[j], v[j][4], v[j][5], R, gamma, f, aq, err, chk);
          } else {
            face_terms<UP>(v[j][0], v[j][1], v[j][2], v[j][3], v[j][4], v[j][5], gamma, f, aq, al, err, chk);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
          if constexpr (UP) {
            if constexpr (VAR == 0) {
              // own points keep α only: α·q is formed again from the stage
              // input held in registers (the same products)
              if (OWNQ && it < RY) {
                cache[it][0] = al;
              } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) cache[it][q] = aq[q];
              }
            } else if constexpr (VAR == 1) {
              cache[it][0] = al;
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[(4 + q) * NP + l] = aq[q];
            }
          }
        }
      }
    }
    __syncthreads();
    // ξ-stencil of one LDS slot group into acc with sign sg
    // acc = sg·idx_ξ·Σ (first pass) or acc += sg·idx_ξ·Σ
    auto xi_pass = [&](auto blk, int slot, CT sg, bool first) {
      using BB = decltype(blk);
#pragma unroll
      for (int s = 0; s < RY; ++s) {
        if (!(yin[s] && any)) continue;
        const CT* line = lds + (size_t)(tr * RY + s) * WX;
        if constexpr (EDGE) {
          if (!interior) {
            // closure rows: each coefficient read once for the four components
#pragma unroll
            for (int i = 0; i < RX; ++i) {
              CT o[4] = {CT(0), CT(0), CT(0), CT(0)};
              if (xin[i])
                row_sum4_tab<BB>(line + slot * NP, NP, 1, xb + i, nx, rx0, tabx + BB::table * OX::M * OX::N,
                                 tabx + NK * OX::M * OX::N + BB::table * LPX, o);
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[s][i][q] = first ? sg * o[q] : fma_t(sg, o[q], acc[s][i][q]);
            }
            continue;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const CT* lq = line + (slot + q) * NP;
          // the window covers the ξ halo: the edge instantiation took the
          // closure path above on tiles without it, and the interior
          // instantiation only gets tiles off the block borders (x0 >= TX >
          // HX, so the region starts at x0 - HX; rows it writes read at most
          // HX columns past them, inside the block and the region)
          CT o[RX];
          row_window<BB, RX>(lq + (xb - rx0 - HX), o);
#pragma unroll
          for (int i = 0; i < RX; ++i) acc[s][i][q] = first ? sg * o[i] : fma_t(sg, o[i], acc[s][i][q]);
        }
      }
    };
    xi_pass(BD{}, 0, CT(-idx_xi), true);
    if constexpr (UP) {
      using BS = Blk<OX, SBP_DISS>;
      if constexpr (VAR != 2) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITXE; ++it) {
          int r, c;
          if (xi_pt(it, r, c)) {
            const int l = r * WX + c;
            if constexpr (NC == 4) {
              if (OWNQ && it < RY) {
#pragma unroll
                for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * ownq[OWNQ ? it : 0][q];
              } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][q];
              }
            } else {
              double qq[4];
              load_q((uint32_t)((y0 + r) * ld + rx0 + c), qq);
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * qq[q];
            }
          }
        }
        __syncthreads();
        xi_pass(BS{}, 0, CT(idx_xi), false);
      } else {
        xi_pass(BS{}, 4, CT(idx_xi), false);
      }
    }
    __syncthreads();
  }

  // Finish operands at the output points, all loaded before any store.
  // FP64 (see the finish), with q the stage input held from the η phase:
  //   S0 q = y_n;  S1 fa = y_n, q = Y2;  S2 fa = y_n, q = Y3;  S3 fa = W, q = Y4
  //   BS0 q = y_n;  BS1 fa = y_n, q = Y2;  BS2, BS3 fa = y_n, fb = Y2, q = Y3 / the solution
  // mixed: S3 fd = Δ_1, Δ_2 (S0-S2 need no operand)
  double fJ[RY][RX], fa[RY][RX][4], fb[RY][RX][4];
  f4 fd[MIX ? RY : 1][MIX ? RX : 1][2];
  static_assert(OWNQ, "the stage combination takes the stage input from the η phase");
  auto load_finish = [&]() {
#pragma unroll
    for (int s = 0; s < RY; ++s)
#pragma unroll
      for (int i = 0; i < RX; ++i) {
        const uint32_t o = (xin[i] && yin[s]) ? (uint32_t)((yb + s) * ld + xb + i) : (uint32_t)(y0 * ld + x0);
        if constexpr (OWNQ) {
          fJ[s][i] = jsave[s * NT + tid];
        } else {
          fJ[s][i] = ldg(met, o);
        }
        if (mode == MODE_RHS) continue;
        if constexpr (MIX) {
          if (mode == MODE_S3) {
#pragma unroll
            for (int t = 0; t < 2; ++t) fd[MIX ? s : 0][MIX ? i : 0][t] = ldg_nt(delta_set(B, t), o);
          }
        } else if (ADAPT) {
          if (mode >= MODE_BS1) {
            const double* yn = A.swap ? B.buf[BUF_ACC] : B.buf[BUF_YN];
#pragma unroll
            for (int q = 0; q < 4; ++q) fa[s][i][q] = ldg_nt(yn + q * ps, o);
            if (mode >= MODE_BS2) load_pm_nt(B.buf[BUF_YA], o, fb[s][i]);
          }
        } else if (mode == MODE_S3) {
          load_pm_nt(B.buf[BUF_ACC], o, fa[s][i]);  // W, point-major
        } else if (mode != MODE_S0) {
          // S1, S2: fa = y_n (planes)
#pragma unroll
          for (int q = 0; q < 4; ++q) fa[s][i][q] = ldg_nt((const double*)B.buf[BUF_YN] + q * ps, o);
        }
      }
  };

  // ------------------------------------------------------------ η-phase
  {
    using BD = Blk<OY, SBP_DIFF>;
    const int ry0 = max(0, y0 - HY), ry1 = min(ny, y0 + TY + HY);
    const int yh = ry1 - ry0;
    const int xend = min(x0 + TX, nx);
    const bool interior = y0 >= OY::M && y0 + TY <= ny - OY::M;
    MixRef64 R;
    if constexpr (MIX) R = load_mix_ref(mref, 1);
    // region point (row r from ry0, column c from x0) of this thread's
    // iteration it: own output points first (OWNQ), then the halo rows
    const int yendr = min(y0 + TY, ny);
    const int ntop = y0 - ry0, nbot = ry1 - yendr;
    auto eta_pt = [&](int it, int& r, int& c) -> bool {
      if constexpr (OWNQ) {
        if (it < RY) {
          r = yb - ry0 + it;
          c = tc;
          return yb + it < yendr && x0 + c < xend;
        }
        const int h = tid + (it - RY) * NT;
        if (h < ntop * TX) {
          r = h / TX;
          c = h - r * TX;
        } else {
          const int h2 = h - ntop * TX;
          const int rr = h2 / TX;
          r = yendr - ry0 + rr;
          c = h2 - rr * TX;
          if (rr >= nbot) return false;
        }
        return x0 + c < xend;
      } else {
        const int l = tid + it * NT;
        r = l / TX;
        c = l - r * TX;
        return l < G::NPY && r < yh && x0 + c < xend;
      }
    };
    CT cache[ITYE][NC];
#pragma unroll
    for (int i0 = 0; i0 < ITYE; i0 += LBY) {
      double v[LBY][6];
#pragma unroll
      for (int j = 0; j < LBY; ++j) {
        if (i0 + j >= ITYE) continue;
        int r, c;
        const bool ok = i0 + j < ITYE && eta_pt(i0 + j, r, c);
        const uint32_t o = ok ? (uint32_t)((ry0 + r) * ld + x0 + c) : (uint32_t)(ry0 * ld + x0);
        if (OWNQ && i0 + j < RY) {
          // the own points' stage input from registers; their η metrics are
          // loaded again (a cache hit) rather than held through the ξ phase
#pragma unroll
          for (int q = 0; q < 4; ++q) v[j][q] = ownq[OWNQ ? i0 + j : 0][q];
          const d2 me = ldg(meta, o);
          v[j][4] = me.x;
          v[j][5] = me.y;
        } else {
          load_q(o, v[j]);
          const d2 me = ldg(meta, o);
          v[j][4] = me.x;
          v[j][5] = me.y;
        }
      }
#pragma unroll
      for (int j = 0; j < LBY; ++j) {
        const int it = i0 + j;
        int r, c;
        if (it < ITYE && eta_pt(it, r, c)) {
          const int l = r * TX + c;
          CT f[4], aq[4];
          double al = 0.0;
          // (ρ and p were checked in the ξ phase)
          if constexpr (MIX) {
            face_terms_mix64<UP>(v[j], v[j][4], v[j][5], R, gamma, f, aq, err, !OWNQ);
          } else {
            face_terms<UP>(v[j][0], v[j][1], v[j][2], v[j][3], v[j][4], v[j][5], gamma, f, aq, al, err, !OWNQ);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) lds[q * NP + l] = f[q];
          if constexpr (UP) {
            if constexpr (VAR == 0) {
              // own points keep α only: α·q is formed again from the stage
              // input held in registers (the same products)
              if (OWNQ && it < RY) {
                cache[it][0] = al;
              } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) cache[it][q] = aq[q];
              }
            } else if constexpr (VAR == 1) {
              cache[it][0] = al;
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[(4 + q) * NP + l] = aq[q];
            }
          }
        }
      }
    }
    __syncthreads();
    // acc += sg·idx_η·Σ
    auto eta_pass = [&](auto blk, int slot, CT sg) {
      using BB = decltype(blk);
      if (!any) return;
      if constexpr (EDGE) {
        if (!interior) {
#pragma unroll
          for (int s = 0; s < RY; ++s)
#pragma unroll
            for (int i = 0; i < RX; ++i) {
              CT o[4] = {CT(0), CT(0), CT(0), CT(0)};
              if (yin[s] && xin[i])
                row_sum4_tab<BB>(lds + slot * NP + tc * RX + i, NP, TX, yb + s, ny, ry0,
                                 taby + BB::table * OY::M * OY::N, taby + NK * OY::M * OY::N + BB::table * LPY, o);
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[s][i][q] = fma_t(sg, o[q], acc[s][i][q]);
            }
          return;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const CT* col = lds + (slot + q) * NP + tc * RX;
        // the window covers the η halo (as in the ξ pass: closure tiles took
        // the edge path, the interior instantiation's tiles start at
        // y0 >= TY > HY)
        CT o[RY][RX];
        col_window<BB, RY, RX>(col + (yb - ry0 - HY) * TX, TX, o);
#pragma unroll
        for (int s = 0; s < RY; ++s)
#pragma unroll
          for (int i = 0; i < RX; ++i) acc[s][i][q] = fma_t(sg, o[s][i], acc[s][i][q]);
      }
    };
    eta_pass(BD{}, 0, CT(-idx_eta));
    if constexpr (UP) {
      using BS = Blk<OY, SBP_DISS>;
      if constexpr (VAR != 2) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITYE; ++it) {
          int r, c;
          if (eta_pt(it, r, c)) {
            const int l = r * TX + c;
            if constexpr (NC == 4) {
              if (OWNQ && it < RY) {
#pragma unroll
                for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * ownq[OWNQ ? it : 0][q];
              } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][q];
              }
            } else {
              double qq[4];
              load_q((uint32_t)((ry0 + r) * ld + x0 + c), qq);
#pragma unroll
              for (int q = 0; q < 4; ++q) lds[q * NP + l] = cache[it][0] * qq[q];
            }
          }
        }
        __syncthreads();
        eta_pass(BS{}, 0, CT(idx_eta));
      } else {
        eta_pass(BS{}, 4, CT(idx_eta));
      }
    }
  }

  // ------------------------------------------------------------ finish
  load_finish();
  KSum esum = {0.0, 0.0};  // MODE_BS3: this thread's error terms
  // (b_s·dt)·k and (a_s·dt)·k with this block's step (integrate, lib.rs:188-211)
  const double dtb = A.dt * B.dt_scale;
  const double sacc = A.rk_b * dtb, snext = A.rk_a * dtb;
#pragma unroll
  for (int s = 0; s < RY; ++s)
#pragma unroll
  for (int i = 0; i < RX; ++i) {
    if (!(xin[i] && yin[s])) continue;
    const uint32_t o = (uint32_t)((yb + s) * ld + xb + i);
    const double ij = frcp(fJ[s][i]);
    double k[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) k[q] = (double)acc[s][i][q] * ij;
    if constexpr (EDGE) {
      // the characteristic SATs of the block-edge points (from the prologue),
      // summed in the order of SAT_characteristics (north, south, east, west;
      // corners get two)
      const int gx = xb + i, gy = yb + s;
      if (gx == 0 || gx == nx - 1 || gy == 0 || gy == ny - 1) {
        double sat[4] = {0.0, 0.0, 0.0, 0.0};
        auto add = [&](int slot) {
#pragma unroll
          for (int q = 0; q < 4; ++q) sat[q] += satl[4 * slot + q];
        };
        if (gy == ny - 1) add(gx - x0);
        if (gy == 0) add(TX + gx - x0);
        if (gx == nx - 1) add(2 * TX + gy - y0);
        if (gx == 0) add(2 * TX + TY + gy - y0);
#pragma unroll
        for (int q = 0; q < 4; ++q) k[q] += sat[q];
      }
    }
    if (mode == MODE_RHS) {
#pragma unroll
      for (int q = 0; q < 4; ++q) stg_nt((A.rhs_out_sel ? B.rk[A.rhs_out_sel - 1] : B.rhs_out) + q * ps, o, k[q]);
    } else if constexpr (MIX) {
      // Mixed stage combination: stages 1-3 store only their FP32 increment
      // Δ_{s+1} = a_s·dt·k_s (the next stage input is y_n + Δ_{s+1}); the
      // last stage sums the increments in FP64,
      // y_{n+1} = y_n + (Δ_1 + 2Δ_2 + Δ_3)/3 + dt/6·k_4, the classical RK4
      // sum with dt·k_1 = 2Δ_1, dt·k_2 = 2Δ_2, dt·k_3 = Δ_3 (the increments
      // the stage inputs used), formed from the stage input y_n + Δ_3 held
      // since the ξ phase as (y_n + Δ_3) + (Δ_1 + 2Δ_2 - 2Δ_3)/3 + dt/6·k_4
      // and written into the other FP64 buffer (the roles swap after the
      // step).
      if (mode == MODE_S3) {
        double* yo = acc_buf(B, A);
        const f4(&d)[2] = fd[MIX ? s : 0][MIX ? i : 0];
        const f4 d3 = ownd[MIX ? s : 0];
        const double* qs = ownq[OWNQ ? s : 0];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double inc = ((double)d[0][q] + 2.0 * ((double)d[1][q] - (double)d3[q])) * (1.0 / 3.0);
          stg_nt(yo + q * ps, o, qs[q] + (inc + sacc * k[q]));
        }
      } else {
        f4 dv;
#pragma unroll
        for (int q = 0; q < 4; ++q) dv[q] = (float)(snext * k[q]);
        stg_nt(delta_out(B, mode), o, dv);
      }
    } else {
      // FP64 stage combination, classical RK4 (utils/integrate/src/lib.rs:30-35,176-222):
      // Y2 = y + dt/2 k1, Y3 = y + dt/2 k2, Y4 = y + dt k3 and
      // y_{n+1} = y + dt/6 (k1 + 2k2 + 2k3 + k4) = (Y2 + 2 Y3 + Y4 - y)/3 + dt/6 k4,
      // an identity of the tableau (dt k1 = 2(Y2 - y), dt k2 = 2(Y3 - y),
      // dt k3 = Y4 - y). Stage 2, which holds its input Y2 since the ξ phase
      // and reads y to form Y3, also writes W = Y2 + 2 Y3 - y (into ACC), so
      // that stage 3 reads only y and stage 4 one operand besides its input:
      // y_{n+1} = (W + Y4)/3 + dt/6 k4. Y4 goes over Y2, which no stage reads
      // after stage 2. No accumulator is kept between stages; the rounding
      // differs from the reference's running sum by a few ulps of y.
      const double* qs = ownq[OWNQ ? s : 0];  // the stage input here
      // Y2, Y3, Y4 and W point-major, y_{n+1} in planes
      const double* fs = fa[s][i];
      if constexpr (ADAPT) {
        // Embedded Bogacki-Shampine trial step (integrate/src/lib.rs:138-151,
        // 176-251) with step g: Y2 = y + g/2 k0, Y3 = y + 3g/4 k1, the
        // solution f = y + g (2/9 k0 + 1/3 k1 + 4/9 k2) and the embedded
        // z = y + g (7/24 k0 + 1/4 k1 + 1/3 k2 + 1/8 k3). With the tableau's
        // identities g k0 = 2 (Y2 - y), g k1 = 4/3 (Y3 - y):
        //   f = y + 4/9 ((Y2 - y) + (Y3 - y)) + 4/9 g k2,
        //   z - y = 1/4 (Y2 - y) + 3/4 (f - y) + g/8 k3,
        // so BS3 needs y, Y2 and its own input f, and z is never stored: its
        // H-norm distance from y (advance_adaptive's error, lib.rs:170-176)
        // is summed per tile. (snext: g/2, 3g/4, 4g/9; sacc: g/8.)
        const double* f2 = fb[s][i];
        if (mode == MODE_BS0) {
          store_pm_nt(B.buf[BUF_YA], o, qs[0] + snext * k[0], qs[1] + snext * k[1], qs[2] + snext * k[2],
                      qs[3] + snext * k[3]);
        } else if (mode == MODE_BS1) {
          store_pm_nt(B.buf[BUF_YB], o, fs[0] + snext * k[0], fs[1] + snext * k[1], fs[2] + snext * k[2],
                      fs[3] + snext * k[3]);
        } else if (mode == MODE_BS2) {
          double* yo = A.swap ? B.buf[BUF_YN] : B.buf[BUF_ACC];
#pragma unroll
          for (int q = 0; q < 4; ++q)
            stg_nt(yo + q * ps, o, fs[q] + ((4.0 / 9.0) * ((f2[q] - fs[q]) + (qs[q] - fs[q])) + snext * k[q]));
        } else {
          const double w = hweight<OX, OY>(xb + i, nx, yb + s, ny);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double d = 0.25 * (f2[q] - fs[q]) + 0.75 * (qs[q] - fs[q]) + sacc * k[q];
            ksum_add(esum, d * d * w);
          }
        }
      } else if (mode == MODE_S0) {
        store_pm_nt(B.buf[BUF_YA], o, qs[0] + snext * k[0], qs[1] + snext * k[1], qs[2] + snext * k[2],
                    qs[3] + snext * k[3]);
      } else if (mode == MODE_S1) {
        double y3[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) y3[q] = fs[q] + snext * k[q];
        store_pm_nt(B.buf[BUF_YB], o, y3[0], y3[1], y3[2], y3[3]);
        store_pm_nt(B.buf[BUF_ACC], o, fma_t(2.0, y3[0], qs[0]) - fs[0], fma_t(2.0, y3[1], qs[1]) - fs[1],
                    fma_t(2.0, y3[2], qs[2]) - fs[2], fma_t(2.0, y3[3], qs[3]) - fs[3]);
      } else if (mode == MODE_S2) {
        store_pm_nt(B.buf[BUF_YA], o, fs[0] + snext * k[0], fs[1] + snext * k[1], fs[2] + snext * k[2],
                    fs[3] + snext * k[3]);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) stg_nt(B.buf[BUF_YN] + q * ps, o, (fa[s][i][q] + qs[q]) * (1.0 / 3.0) + sacc * k[q]);
      }
    }
  }
  if constexpr (ADAPT) {
    if (mode == MODE_BS3) {
      // the tile's sum of the adaptive step's error terms, compensated
      __shared__ KSum wsum[NT / 64];
      esum = wave_ksum(esum);
      if ((tid & 63) == 0) wsum[tid >> 6] = esum;
      __syncthreads();
      if (tid == 0) {
        KSum t = wsum[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) ksum_merge(t, wsum[w]);
        const int64_t slot = B.tile_off + (int64_t)T.ty * B.ntx + T.tx;
        A.partials[2 * slot] = t.s;
        A.partials[2 * slot + 1] = t.c;
      }
    }
  }
  if (err) record_error(A.error, err);
  // the time update of an RK4 step (StageArgs::time_w), after the tile's
  // own work so that the store orders none of it
  if (A.time_w && blockIdx.x == 0 && tid == 0) {
    const double t = *A.time;
    *A.time_w = mode == MODE_S0 ? t : t + A.dt;
  }
}

}  // namespace sbp
