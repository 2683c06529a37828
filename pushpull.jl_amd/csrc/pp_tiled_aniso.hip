// pp_tiled_aniso.hip -- the tiled kernels for degrees that differ between axes
// (src/pushpull_general.jl:27-30 allows a degree per axis).
//
// A triple whose degrees differ runs on the 16x8x8 kernels compiled for the largest (DP
// taps a window), each axis's own dp taps placed in that window at their shift and the
// window's other taps weighted 0 (pp_tiled_impl.hpp: weights_win). Boxes, staging, the
// fixed-point scatters and the flush are those of the isotropic kernels. Where a zero
// weight could meet a non-finite image value (the reference never reads that tap), the
// staged box is cleared of non-finite values and its channel's samples take the global
// path, which evaluates every axis with its own taps only (with_axes). Compiled in a
// translation unit of its own, so that the isotropic kernels keep their code and register
// budgets and both build in parallel.
#include "pp_tiled_impl.hpp"

namespace pp {

template <int DP>
static hipError_t launch_aniso_deg(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  const int cap = cfg.cap > 0 ? cfg.cap : default_cap(op, DP);
  return launch_shape<DP, 16, 8, 8, true>(op, a, nbatch, s, cap);
}

hipError_t launch_tiled_aniso(Op op, const KArgs& a, int nbatch, hipStream_t s, TileCfg cfg) {
  if (std::min(a.dpx, std::min(a.dpy, a.dpz)) < 1) return hipErrorNotSupported;
  switch (std::max(a.dpx, std::max(a.dpy, a.dpz))) {
    case 2: return launch_aniso_deg<2>(op, a, nbatch, s, cfg);
    case 3: return launch_aniso_deg<3>(op, a, nbatch, s, cfg);
    case 4: return launch_aniso_deg<4>(op, a, nbatch, s, cfg);
    case 5: return launch_aniso_deg<5>(op, a, nbatch, s, cfg);
    default: return hipErrorNotSupported;
  }
}

const void* batch_kernel_aniso(Op op, int dp) { return batch_kernel_for<true>(op, dp); }

}  // namespace pp
