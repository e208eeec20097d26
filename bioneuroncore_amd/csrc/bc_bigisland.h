// bc_bigisland.hip — step 2.5 for islands larger than one tile (more than 256 segments or 256
// connections): one 1024-thread workgroup per island, every per-segment and per-connection
// quantity in global memory (the island's slice of BigScratch), threads striding over the
// island's segments and connections with a workgroup barrier between phases and sweeps.
//
// The arithmetic is the tile kernel's (k_diffusion), operation for operation: per-segment sums
// over the segment's incidence list in global connection order (reference
// include/diffusion_kernel.h:41-104 for phi_e, :151-238 intracellular, :245-281 extracellular,
// :284-345 Fick, :349-363 bath), the same correctly rounded division helpers, the same exact
// early end of the fixed-count sweeps (bitwise fixed point or two-cycle of the island's
// iterate), residual-terminated sweeps and cable mode.
#pragma once
#include <hip/hip_runtime.h>
#include "bc_device.h"
#include "bc_state.h"

namespace bc {

constexpr int kBigThreads = 1024;

// The big islands: island k holds device segments [seg0[k], seg0[k] + nseg[k]) and device
// connections [conn0[k], conn0[k] + nconn[k]); big islands follow all tiles in device order,
// so (segment - seg_base) and (connection - conn_base) index the scratch planes below.
struct BigSet {
    const uint32_t* seg0;
    const uint32_t* nseg;
    const uint32_t* conn0;
    const uint32_t* nconn;
    const uint32_t* ca;   // per big connection: device index of end a
    const uint32_t* cb;   //   and of end b
    uint32_t count;
    uint32_t seg_base, conn_base;
    uint32_t nseg_all, nconn_all;   // plane strides
    // per segment
    double* phi;    // [2][nseg_all]
    double* diag;
    float* vf;      // [2][nseg_all]
    float* base;
    float* den;
    double* dacc;   // [7][nseg_all]
    float* dsp;     // [33][nseg_all]: 15 transmitter, 15 intra transmitter, ATP, temp, txn deltas
    // per connection
    double* gphi;
    float* cac;
    float* cdist;
    float* flux;    // [7][nconn_all] intracellular
    float* cg;
    float* cterm;
    double* eflux;  // [7][nconn_all] extracellular
};

template <bool kRes>
__global__ void k_diffusion_big(Net d, BigSet bs, float dt);

}  // namespace bc
