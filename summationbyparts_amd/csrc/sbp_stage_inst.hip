// Instantiations of the fused stage kernel for one ξ-operator (SBP_OPX, set
// by the build) and every η-operator, RHS kind and tile shape.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sbp_kernels.hpp"
#include "sbp_stage.hpp"
#include "sbp_stage.hpp"

#ifndef SBP_OPX
#error "SBP_OPX (operator type) and SBP_OPX_ID must be defined by the build"
#endif

#define SBP_CAT2(a, b) a##b
#define SBP_CAT(a, b) SBP_CAT2(a, b)

namespace sbp {

// Kernel configurations (tile TX x TY, RY points per thread, caching
// variant); the list matches kTileShapes in sbp_misc.hip.
// The mixed form always keeps flux and α·q together in LDS (VAR 2).
// The embedded Bogacki-Shampine modes (FP64) run in their own instantiation.
template <class OX, class OY, bool UP, bool MIX, bool EDGE, int TX, int TY, int RX, int RY, int VAR>
static void launch_one(int mode, const StageArgs& args, int ntiles, hipStream_t stream) {
  if constexpr (MIX) {
    stage_kernel<OX, OY, UP, TX, TY, RX, RY, 2, true, EDGE, false>
        <<<ntiles, (TX / RX) * (TY / RY), 0, stream>>>(args, mode);
  } else if (mode >= MODE_BS0) {
    stage_kernel<OX, OY, UP, TX, TY, RX, RY, VAR, false, EDGE, true>
        <<<ntiles, (TX / RX) * (TY / RY), 0, stream>>>(args, mode);
  } else {
    stage_kernel<OX, OY, UP, TX, TY, RX, RY, VAR, false, EDGE, false>
        <<<ntiles, (TX / RX) * (TY / RY), 0, stream>>>(args, mode);
  }
}

// Every configuration for equal operator pairs; pairs of different ξ and η
// operators get the three shapes the autotuner picks on the example
// workloads (32x16 in both variants, 16x16), which keeps the library's size
// and build time in bounds.
template <class OX, class OY, bool UP, bool MIX, bool EDGE>
static hipError_t launch_shape(int cfg, int mode, const StageArgs& args, int ntiles, hipStream_t stream) {
  constexpr bool full = std::is_same<OX, OY>::value;
  switch (cfg) {
    case 0: launch_one<OX, OY, UP, MIX, EDGE, 32, 16, 1, 2, 1>(mode, args, ntiles, stream); break;
    case 1: launch_one<OX, OY, UP, MIX, EDGE, 32, 16, 1, 2, 0>(mode, args, ntiles, stream); break;
    case 7: launch_one<OX, OY, UP, MIX, EDGE, 16, 16, 1, 2, 0>(mode, args, ntiles, stream); break;
#define SBP_CFG(N, TX, TY, RX, RY, VAR)                                                            \
  case N:                                                                                        \
    if constexpr (full) {                                                                        \
      launch_one<OX, OY, UP, MIX, EDGE, TX, TY, RX, RY, VAR>(mode, args, ntiles, stream);        \
      break;                                                                                     \
    }                                                                                            \
    return hipErrorNotSupported;
    SBP_CFG(2, 32, 24, 1, 2, 0)
    SBP_CFG(3, 32, 20, 1, 2, 0)
    SBP_CFG(4, 32, 12, 1, 2, 0)
    SBP_CFG(5, 64, 16, 1, 2, 0)
    SBP_CFG(6, 32, 24, 1, 2, 1)
    SBP_CFG(8, 16, 24, 1, 2, 0)
    SBP_CFG(9, 32, 8, 1, 1, 0)
    SBP_CFG(10, 32, 16, 1, 1, 0)
    SBP_CFG(11, 64, 8, 1, 1, 0)
    SBP_CFG(12, 32, 12, 1, 1, 0)
#undef SBP_CFG
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class OX, class OY, bool MIX, bool EDGE>
static hipError_t launch_pair_p(bool up, int shape, int mode, const StageArgs& args, int ntiles, hipStream_t stream) {
  if (up) {
    if constexpr (OX::is_upwind && OY::is_upwind)
      return launch_shape<OX, OY, true, MIX, EDGE>(shape, mode, args, ntiles, stream);
    return hipErrorInvalidValue;
  }
  return launch_shape<OX, OY, false, MIX, EDGE>(shape, mode, args, ntiles, stream);
}

// edge: the tiles on block edges, which add the edge SATs
template <class OX, class OY>
static hipError_t launch_pair(bool up, bool edge, int shape, int mode, const StageArgs& args, int ntiles,
                              hipStream_t stream) {
  if (args.mixed) {
    if (edge) return launch_pair_p<OX, OY, true, true>(up, shape, mode, args, ntiles, stream);
    return launch_pair_p<OX, OY, true, false>(up, shape, mode, args, ntiles, stream);
  }
  if (edge) return launch_pair_p<OX, OY, false, true>(up, shape, mode, args, ntiles, stream);
  return launch_pair_p<OX, OY, false, false>(up, shape, mode, args, ntiles, stream);
}

hipError_t SBP_CAT(launch_stage_opxi_, SBP_OPX_ID)(int op_eta, bool up, bool edge, int shape, int mode, const StageArgs& args,
                                                   int ntiles, hipStream_t stream) {
  using OX = ops::SBP_OPX;
  switch (op_eta) {
    case UPWIND4: return launch_pair<OX, ops::Upwind4>(up, edge, shape, mode, args, ntiles, stream);
    case UPWIND4H2: return launch_pair<OX, ops::Upwind4h2>(up, edge, shape, mode, args, ntiles, stream);
    case UPWIND9: return launch_pair<OX, ops::Upwind9>(up, edge, shape, mode, args, ntiles, stream);
    case UPWIND9H2: return launch_pair<OX, ops::Upwind9h2>(up, edge, shape, mode, args, ntiles, stream);
    case SBP4: return launch_pair<OX, ops::SBP4>(up, edge, shape, mode, args, ntiles, stream);
    case SBP8: return launch_pair<OX, ops::SBP8>(up, edge, shape, mode, args, ntiles, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sbp
