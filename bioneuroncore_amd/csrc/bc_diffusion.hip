// bc_diffusion.hip — step 2.5 kernels of the 64/128/256-segment tiles (fixed sweeps, residual-
// terminated sweeps, and the cable mode that records the Jacobi iterates for the adjoint).
#include "bc_tick.h"

namespace bc {

// 64-lane tiles with fixed sweeps run at 5 wavefronts per SIMD (at most 102 VGPRs; LDS: 5
// workgroups of four 8 KB tiles fill a CU's 160 KB): the sweeps are latency chains, and the fifth
// wavefront hides more of them (0.97 ms against 1.02 ms at four, 3.2M segments). The residual
// variant's island flags do not leave LDS for a fifth.
template <int TS, bool kRes, bool kRec>
__global__ __launch_bounds__(TS == 64 ? 64 * kDiffMaxTiles64 : TS)
__attribute__((amdgpu_waves_per_eu(TS == 64 && !kRes && !kRec ? 5 : 1))) void k_diffusion(Net d, TileSet ts, float dt) {
    diffusion_body<TS, kRes, kRec>(d, ts, dt);
}

template __global__ void k_diffusion<64, false, false>(Net, TileSet, float);
template __global__ void k_diffusion<128, false, false>(Net, TileSet, float);
template __global__ void k_diffusion<256, false, false>(Net, TileSet, float);
template __global__ void k_diffusion<64, true, false>(Net, TileSet, float);
template __global__ void k_diffusion<128, true, false>(Net, TileSet, float);
template __global__ void k_diffusion<256, true, false>(Net, TileSet, float);
template __global__ void k_diffusion<64, false, true>(Net, TileSet, float);
template __global__ void k_diffusion<128, false, true>(Net, TileSet, float);
template __global__ void k_diffusion<256, false, true>(Net, TileSet, float);

}  // namespace bc
