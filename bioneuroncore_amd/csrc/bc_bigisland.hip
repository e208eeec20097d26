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
#include "bc_bigisland.h"

namespace bc {

// Workgroup reductions over the island (barriers included).
__device__ static inline bool big_any(bool p) { return __syncthreads_or(p) != 0; }
__device__ static inline double big_max(double v, unsigned long long* slot) {
    if (threadIdx.x == 0) *slot = 0;
    __syncthreads();
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));   // v >= 0: bits order as values
    __syncthreads();
    const double r = __longlong_as_double((long long)*slot);
    __syncthreads();
    return r;
}

template <bool kRes>
__global__ __launch_bounds__(kBigThreads) void k_diffusion_big(Net d, BigSet bs, float dt) {
    __shared__ unsigned long long red;
    const uint32_t isl = blockIdx.x;
    if (isl >= bs.count) return;
    const uint32_t s0 = bs.seg0[isl], ns = bs.nseg[isl], c0 = bs.conn0[isl], nc = bs.nconn[isl];
    const size_t P = d.npad, SN = bs.nseg_all, CN = bs.nconn_all;
    const uint32_t T = blockDim.x, t = threadIdx.x;
    auto si = [&](uint32_t g) { return (size_t)(g - bs.seg_base); };   // scratch index of segment g
    auto ci = [&](uint32_t c) { return (size_t)(c - bs.conn_base); };
    auto vol_of = [&](uint32_t g) {
        float v = (float)d.vol[g];
        return v < 1e-6f ? 1e-6f : v;
    };
    // a connection's other end as seen from segment g's incidence entry
    auto other = [&](uint32_t en) {
        const uint32_t c = c0 + (en >> 1);
        return (en & 1) ? bs.ca[ci(c)] : bs.cb[ci(c)];
    };
    const int sweeps = kRes ? d.jac_max : 50;

    // ---- connection geometry (diffusion_kernel.h:72-80,162-172) ----
    for (uint32_t k = t; k < nc; k += T) {
        const uint32_t c = c0 + k;
        const uint32_t a = bs.ca[ci(c)], b = bs.cb[ci(c)];
        float ra = d.radius[a], rb = d.radius[b];
        float rmin = ra < rb ? ra : rb;
        float cac = kPi * rmin * rmin;
        if (cac < 1e-6f) cac = 1e-6f;
        float cdist = (d.length[a] + d.length[b]) * 0.5f;
        if (cdist < 0.1f) cdist = 0.1f;
        bs.cac[ci(c)] = cac;
        bs.cdist[ci(c)] = cdist;
    }
    // ---- phi_e (diffusion_kernel.h:41-104) ----
    if (!d.cable) {
        for (uint32_t k = t; k < ns; k += T) {   // sigma_e in diag for now, phi_0 = 0
            const uint32_t g = s0 + k;
            double sg = 0.0;
#pragma unroll
            for (int q = 0; q < 7; q++) sg += kZ[q] * kZ[q] * kDiff[q] * d.ex[q * P + g];
            bs.diag[si(g)] = kSigmaE * sg;
            bs.phi[si(g)] = 0.0;
        }
        __syncthreads();
        for (uint32_t k = t; k < nc; k += T) {
            const uint32_t c = c0 + k;
            const uint32_t a = bs.ca[ci(c)], b = bs.cb[ci(c)];
            bs.gphi[ci(c)] = 0.5 * (bs.diag[si(a)] + bs.diag[si(b)]) * bs.cac[ci(c)] / bs.cdist[ci(c)];
        }
        __syncthreads();
        for (uint32_t k = t; k < ns; k += T) {   // diag = sigma * area + sum of g, in connection order
            const uint32_t g = s0 + k;
            double dg = bs.diag[si(g)] * d.area[g] / 1.0f;
            for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) dg += bs.gphi[ci(c0 + (d.csr_ent[e] >> 1))];
            bs.diag[si(g)] = dg;
        }
        __syncthreads();
        int cur = 0;
        for (int it = 0; it < sweeps; it++) {
            bool d1 = false, d2 = false;
            double res = 0.0;
            for (uint32_t k = t; k < ns; k += T) {
                const uint32_t g = s0 + k;
                double pn = 0.0;
                for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) {
                    const uint32_t en = d.csr_ent[e];
                    const double gg = bs.gphi[ci(c0 + (en >> 1))];
                    if (gg <= 0.0) continue;
                    pn += gg * bs.phi[cur * SN + si(other(en))];
                }
                const double diag = bs.diag[si(g)];
                double ph = (diag > 1e-9) ? div_rcp(pn - d.ibuf[g], diag, rcp_prep(diag)) : 0.0;
                if (ph > 10.0) ph = 10.0;
                if (ph < -10.0) ph = -10.0;
                const double old = bs.phi[cur * SN + si(g)];
                // plane cur ^ 1 holds x_{j-2} until this segment's own write (other segments
                // read plane cur only); on the first sweep it is not read for the test
                const double two = bs.phi[(cur ^ 1) * SN + si(g)];
                bs.phi[(cur ^ 1) * SN + si(g)] = ph;
                if (kRes) {
                    const double r = fabs(ph - old);
                    res = r > res ? r : res;
                } else {
                    d1 |= __double_as_longlong(ph) != __double_as_longlong(old);
                    d2 |= __double_as_longlong(ph) != __double_as_longlong(two);
                }
            }
            cur ^= 1;
            if (kRes) {
                if (big_max(res, &red) <= d.jac_tol) break;
            } else {
                const bool a1 = big_any(d1), a2 = big_any(d2);
                if (!a1) break;                        // fixed point: phi[cur] holds x_j
                if (it >= 1 && !a2) {                  // two-cycle: x_j, x_{j-1} alternate
                    if ((50 - (it + 1)) & 1) cur ^= 1;
                    break;
                }
            }
        }
        if (cur) {   // the final iterate into plane 0
            __syncthreads();
            for (uint32_t k = t; k < ns; k += T) bs.phi[si(s0 + k)] = bs.phi[SN + si(s0 + k)];
        }
    } else {
        for (uint32_t k = t; k < ns; k += T) bs.phi[si(s0 + k)] = 0.0;
    }
    __syncthreads();

    // ---- intracellular Nernst-Planck, 5 substeps (diffusion_kernel.h:151-238) ----
    const int kSub = 5;
    const float dts = dt / kSub;
    const double rdts = 1.0 / (double)dts;
    for (int sub = 0; sub < kSub; sub++) {
        for (uint32_t k = t; k < nc; k += T) {
            const uint32_t c = c0 + k;
            if (!d.conn_intra[c]) continue;
            const uint32_t a = bs.ca[ci(c)], b = bs.cb[ci(c)];
            const float cac = bs.cac[ci(c)], cdist = bs.cdist[ci(c)];
            float sgi = 0.0f;
#pragma unroll
            for (int q = 0; q < 7; q++) {
                float cbar = 0.5f * (d.in[q * P + a] + d.in[q * P + b]);
                sgi += kZ[q] * kZ[q] * kDiff[q] * cbar;
            }
            float gax = div_wide(kSigmaE * sgi * cac, cdist, 1.0 / (double)cdist);
            const float csc = d.conn_scale ? d.conn_scale[c] : 1.0f;
            if (d.conn_scale) gax *= csc;
            bs.cg[ci(c)] = gax;
            float dV = (d.V[a] + bs.phi[si(a)]) - (d.V[b] + bs.phi[si(b)]);
            float inp = 0.0f;
#pragma unroll
            for (int q = 0; q < 7; q++) {
                float ca = d.in[q * P + a], cb = d.in[q * P + b];
                float x = kZ[q] * kFrtMv * dV;
                float kin = kDiff[q] * cac * dts / cdist;
                float fl = kin * (ca * bernoulli(-x, ExpTabGlobal{}) - cb * bernoulli(x, ExpTabGlobal{}));
                if (d.conn_scale) fl *= csc;
                bs.flux[q * CN + ci(c)] = fl;
                inp += div_wide(kZ[q] * kChargeToPa * fl, dts, rdts);
            }
            bs.cterm[ci(c)] = inp - gax * dV;
        }
        __syncthreads();
        for (uint32_t k = t; k < ns; k += T) {   // per segment, in connection order
            const uint32_t g = s0 + k;
            const float volf = vol_of(g);
            const double rvol = 1.0 / (double)volf;
            float gdiag = 0.0f, inpd = 0.0f;
            double dacc[7];
#pragma unroll
            for (int q = 0; q < 7; q++) dacc[q] = 0.0;
            for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) {
                const uint32_t en = d.csr_ent[e];
                const uint32_t c = c0 + (en >> 1);
                if (!d.conn_intra[c]) continue;
                gdiag += bs.cg[ci(c)];
                if (en & 1) {
                    inpd -= bs.cterm[ci(c)];
#pragma unroll
                    for (int q = 0; q < 7; q++) dacc[q] += div_f24((double)bs.flux[q * CN + ci(c)], (double)volf, rvol);
                } else {
                    inpd += bs.cterm[ci(c)];
#pragma unroll
                    for (int q = 0; q < 7; q++) dacc[q] -= div_f24((double)bs.flux[q * CN + ci(c)], (double)volf, rvol);
                }
            }
#pragma unroll
            for (int q = 0; q < 7; q++) bs.dacc[q * SN + si(g)] = dacc[q];
            float cdt = d.cap[g] / dts;
            if (cdt < 1e-9f) cdt = 1e-9f;
            const float vold = (float)d.V[g];
            bs.vf[si(g)] = vold;
            bs.den[si(g)] = cdt + gdiag;
            bs.base[si(g)] = cdt * vold + inpd;
        }
        __syncthreads();
        // Jacobi sweeps of V (x_0 = V before the substep), with the exact early end of the phi
        // sweeps above
        int vc = 0;
        for (int it = 0; it < sweeps; it++) {
            bool d1 = false, d2 = false;
            double res = 0.0;
            for (uint32_t k = t; k < ns; k += T) {
                const uint32_t g = s0 + k;
                float off = 0.0f;
                for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) {
                    const uint32_t en = d.csr_ent[e];
                    const uint32_t c = c0 + (en >> 1);
                    if (!d.conn_intra[c]) continue;
                    off += bs.cg[ci(c)] * bs.vf[vc * SN + si(other(en))];
                }
                const float den = bs.den[si(g)];
                const float v = div_wide(bs.base[si(g)] + off, den, 1.0 / (double)den);
                const float old = bs.vf[vc * SN + si(g)];
                const float two = bs.vf[(vc ^ 1) * SN + si(g)];   // x_{j-2} (x_0 on the first sweep)
                bs.vf[(vc ^ 1) * SN + si(g)] = v;
                if (kRes) {
                    const double r = (double)fabsf(v - old);
                    res = r > res ? r : res;
                } else {
                    d1 |= __float_as_uint(v) != __float_as_uint(old);
                    d2 |= __float_as_uint(v) != __float_as_uint(two);
                }
            }
            vc ^= 1;
            if (kRes) {
                if (big_max(res, &red) <= d.jac_tol) break;
            } else {
                const bool a1 = big_any(d1), a2 = big_any(d2);
                if (!a1) break;
                if (it >= 1 && !a2) {
                    if ((50 - (it + 1)) & 1) vc ^= 1;
                    break;
                }
            }
        }
        __syncthreads();
        for (uint32_t k = t; k < ns; k += T) {
            const uint32_t g = s0 + k;
            d.V[g] = bs.vf[vc * SN + si(g)];
            if (!d.cable)
#pragma unroll
                for (int q = 0; q < 7; q++) d.in[q * P + g] += (float)bs.dacc[q * SN + si(g)];
        }
        __syncthreads();
    }
    if (d.cable) return;   // concentrations, ATP, temperature and transmitters are held

    // ---- extracellular Nernst-Planck, 5 substeps (diffusion_kernel.h:245-281) ----
    for (int sub = 0; sub < kSub; sub++) {
        for (uint32_t k = t; k < nc; k += T) {
            const uint32_t c = c0 + k;
            const uint32_t a = bs.ca[ci(c)], b = bs.cb[ci(c)];
            const float cac = bs.cac[ci(c)], cdist = bs.cdist[ci(c)];
            const double dphi = bs.phi[si(a)] - bs.phi[si(b)];
#pragma unroll
            for (int q = 0; q < 7; q++) {
                double ca = d.ex[q * P + a], cb = d.ex[q * P + b];
                float x = kZ[q] * kFrtMv * dphi;
                float kin = kDiff[q] * cac * dts / cdist;
                bs.eflux[q * CN + ci(c)] = kin * (ca * bernoulli(-x, ExpTabGlobal{}) - cb * bernoulli(x, ExpTabGlobal{}));
            }
        }
        __syncthreads();
        for (uint32_t k = t; k < ns; k += T) {
            const uint32_t g = s0 + k;
            const float ve = vol_of(g) * kEvr;
            const double rve = 1.0 / (double)ve;
            double dx[7];
#pragma unroll
            for (int q = 0; q < 7; q++) dx[q] = 0.0;
            for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) {
                const uint32_t en = d.csr_ent[e];
                const uint32_t c = c0 + (en >> 1);
                if (en & 1) {
#pragma unroll
                    for (int q = 0; q < 7; q++) dx[q] += div_f24(bs.eflux[q * CN + ci(c)], (double)ve, rve);
                } else {
#pragma unroll
                    for (int q = 0; q < 7; q++) dx[q] -= div_f24(bs.eflux[q * CN + ci(c)], (double)ve, rve);
                }
            }
#pragma unroll
            for (int q = 0; q < 7; q++) bs.dacc[q * SN + si(g)] = dx[q];
        }
        __syncthreads();
        for (uint32_t k = t; k < ns; k += T) {
            const uint32_t g = s0 + k;
#pragma unroll
            for (int q = 0; q < 7; q++) d.ex[q * P + g] += bs.dacc[q * SN + si(g)];
        }
        __syncthreads();
    }
    // bath relaxation of the extracellular ions (diffusion_kernel.h:349-363) comes after the
    // Fick terms below read nothing of them; done per segment at the end

    // ---- transmitters, ATP, temperature, transcription (diffusion_kernel.h:284-345) ----
    const bool nt_ex = d.nt_ex_live != 0, nt_in = d.nt_in_live != 0;
    for (uint32_t k = t; k < ns; k += T) {
        const uint32_t g = s0 + k;
        const float volf = vol_of(g);
        const double rvol = 1.0 / (double)volf;
        float dnt[15], dnt_in[15];
#pragma unroll
        for (int q = 0; q < 15; q++) { dnt[q] = 0.0f; dnt_in[q] = 0.0f; }
        float datp = 0.0f, dtemp = 0.0f, dtxn = 0.0f;
        for (uint32_t e = d.csr_off[g]; e < d.csr_off[g + 1]; e++) {
            const uint32_t en = d.csr_ent[e];
            const uint32_t c = c0 + (en >> 1);
            const bool role_b = en & 1;
            const uint32_t a = bs.ca[ci(c)], b = bs.cb[ci(c)];
            const float ac = bs.cac[ci(c)], dist = bs.cdist[ci(c)];
            const bool intra = d.conn_intra[c] != 0;
            if (nt_ex)
                for (int q = 0; q < 15; q++) {
                    float dC = d.nt[q * P + a] - d.nt[q * P + b];
                    if (dC < 1e-12f && dC > -1e-12f) continue;
                    float fl = d.nt_diff[q] * ac * dC * dt / dist;
                    if (role_b) dnt[q] += div_wide(fl, volf, rvol);
                    else dnt[q] -= div_wide(fl, volf, rvol);
                }
            if (nt_in && intra)
                for (int q = 0; q < 15; q++) {
                    float dC = d.nt_in[q * P + a] - d.nt_in[q * P + b];
                    if (dC < 1e-12f && dC > -1e-12f) continue;
                    float fl = d.nt_diff[q] * ac * dC * dt / dist;
                    if (role_b) dnt_in[q] += div_wide(fl, volf, rvol);
                    else dnt_in[q] -= div_wide(fl, volf, rvol);
                }
            if (intra) {
                float d1 = d.atp[a] - d.atp[b];
                if (d1 > 1e-12f || d1 < -1e-12f) {
                    float fl = 0.14f * ac * d1 * dt / dist;
                    if (role_b) datp += div_wide(fl, volf, rvol); else datp -= div_wide(fl, volf, rvol);
                }
                float d2 = d.temp[a] - d.temp[b];
                if (d2 > 0.001f || d2 < -0.001f) {
                    float fl = 0.0014f * ac * d2 * dt / dist;
                    if (role_b) dtemp += div_wide(fl, volf, rvol); else dtemp -= div_wide(fl, volf, rvol);
                }
                float d3 = d.txn[a] - d.txn[b];
                if (d3 > 1e-9f || d3 < -1e-9f) {
                    float fl = 0.01f * ac * d3 * dt / dist;
                    if (role_b) dtxn += div_wide(fl, volf, rvol); else dtxn -= div_wide(fl, volf, rvol);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 15; q++) {
            bs.dsp[q * SN + si(g)] = dnt[q];
            bs.dsp[(15 + q) * SN + si(g)] = dnt_in[q];
        }
        bs.dsp[30 * SN + si(g)] = datp;
        bs.dsp[31 * SN + si(g)] = dtemp;
        bs.dsp[32 * SN + si(g)] = dtxn;
    }
    __syncthreads();
    for (uint32_t k = t; k < ns; k += T) {
        const uint32_t g = s0 + k;
#pragma unroll
        for (int q = 0; q < 7; q++) {
            double x = d.ex[q * P + g];
            x += (kBath[q] - x) * dt / 1.0f;
            if (x < 1e-9) x = 1e-9;
            if (x > 1000.0) x = 1000.0;
            d.ex[q * P + g] = x;
        }
        if (nt_in)
            for (int q = 0; q < 15; q++) d.nt_in[q * P + g] += bs.dsp[(15 + q) * SN + si(g)];
        if (nt_ex)
            for (int q = 0; q < 15; q++) d.nt[q * P + g] += bs.dsp[q * SN + si(g)];
        d.atp[g] += bs.dsp[30 * SN + si(g)];
        d.temp[g] += bs.dsp[31 * SN + si(g)];
        d.txn[g] += bs.dsp[32 * SN + si(g)];
    }
}

template __global__ void k_diffusion_big<true>(Net, BigSet, float);
template __global__ void k_diffusion_big<false>(Net, BigSet, float);

}  // namespace bc
