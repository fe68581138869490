// Phase-space projection of tracked beams and beam statistics on the device.
//
// Projection: PhaseSpaceParticle.create_from_running_particle
// (final_code/packages/particles.py:1077-1121) with the same operation order:
//   rel = p - p_ideal; x = XI.rel, y = YI.rel, z = ZI.rel;
//   dv = v - v_ideal; xp = (XI.dv) / sqrt(speed^2 - (XI.dv)^2), yp likewise;
//   delta = (P - P_ideal) / P_ideal with P = speed * relativistic mass.
// Statistics per group of consecutive particles (the beam-size loop of
// cctpy3/run.py:72-103 takes (max - min) / 2 of x and y per group): count,
// min/max of x and y, means, rms about the mean, and the centred second moments
// of delta and of (x, delta) that give the dispersion D = cov(x, delta) / var(delta).
// One workgroup per group; reductions run in a fixed order (deterministic).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace cctamd_detail {
namespace {

constexpr int kBeamWG = 256;

struct Acc {
    double n, mnx, mxx, mny, mxy, sx, sy, sd;
};

__device__ inline Acc acc_zero() {
    const double inf = __builtin_huge_val();
    return Acc{0.0, inf, -inf, inf, -inf, 0.0, 0.0, 0.0};
}

__device__ inline Acc acc_merge(const Acc& a, const Acc& b) {
    return Acc{a.n + b.n, fmin(a.mnx, b.mnx), fmax(a.mxx, b.mxx), fmin(a.mny, b.mny), fmax(a.mxy, b.mxy),
               a.sx + b.sx, a.sy + b.sy, a.sd + b.sd};
}

// Fixed-order tree reduction of one double per thread over the workgroup.
__device__ inline double block_sum(double v, double* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBeamWG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + s];
        __syncthreads();
    }
    const double r = lds[0];
    __syncthreads();
    return r;
}

__device__ inline Acc block_acc(Acc a, Acc* lds) {
    lds[threadIdx.x] = a;
    __syncthreads();
    for (int s = kBeamWG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] = acc_merge(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    const Acc r = lds[0];
    __syncthreads();
    return r;
}

__device__ inline void group_range(const int64_t* __restrict__ group_off, int64_t g, int64_t ptot, int64_t& lo,
                                   int64_t& hi) {
    lo = group_off[g];
    hi = group_off[g + 1];
    lo = lo < 0 ? 0 : (lo > ptot ? ptot : lo);
    hi = hi < lo ? lo : (hi > ptot ? ptot : hi);
}

__device__ inline const double* ideal_of(const double* __restrict__ ideal, const int32_t* __restrict__ group_design,
                                         int64_t g, int64_t ndesign) {
    int64_t d = group_design[g];
    d = d < 0 ? 0 : (d >= ndesign ? ndesign - 1 : d);
    return ideal + 16 * d;
}

#pragma clang fp contract(off)
__global__ __launch_bounds__(kBeamWG) void k_phase_space(const double* __restrict__ ideal, int64_t ndesign,
                                                         const double* __restrict__ state,
                                                         const double* __restrict__ speed,
                                                         const double* __restrict__ momentum, int64_t ptot,
                                                         const int64_t* __restrict__ group_off,
                                                         const int32_t* __restrict__ group_design,
                                                         double* __restrict__ pps, double* __restrict__ stats) {
    __shared__ Acc lacc[kBeamWG];
    __shared__ double lsum[kBeamWG];
    const int64_t g = blockIdx.x;
    int64_t lo, hi;
    group_range(group_off, g, ptot, lo, hi);
    const double* id = ideal_of(ideal, group_design, g, ndesign);
    const double* XI = id + 7;
    const double* YI = id + 10;
    const double* ZI = id + 13;
    Acc a = acc_zero();
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBeamWG) {
        const double r0 = state[i] - id[0], r1 = state[ptot + i] - id[1], r2 = state[2 * ptot + i] - id[2];
        const double x = XI[0] * r0 + XI[1] * r1 + XI[2] * r2;
        const double y = YI[0] * r0 + YI[1] * r1 + YI[2] * r2;
        const double z = ZI[0] * r0 + ZI[1] * r1 + ZI[2] * r2;
        const double v0 = state[3 * ptot + i] - id[3], v1 = state[4 * ptot + i] - id[4],
                     v2 = state[5 * ptot + i] - id[5];
        const double xv = XI[0] * v0 + XI[1] * v1 + XI[2] * v2;
        const double yv = YI[0] * v0 + YI[1] * v1 + YI[2] * v2;
        const double s2 = speed[i] * speed[i];
        const double xp = xv / sqrt(s2 - xv * xv);
        const double yp = yv / sqrt(s2 - yv * yv);
        const double delta = (momentum[i] - id[6]) / id[6];
        pps[i] = x;
        pps[ptot + i] = xp;
        pps[2 * ptot + i] = y;
        pps[3 * ptot + i] = yp;
        pps[4 * ptot + i] = z;
        pps[5 * ptot + i] = delta;
        a.n = a.n + 1.0;
        a.mnx = fmin(a.mnx, x);
        a.mxx = fmax(a.mxx, x);
        a.mny = fmin(a.mny, y);
        a.mxy = fmax(a.mxy, y);
        a.sx = a.sx + x;
        a.sy = a.sy + y;
        a.sd = a.sd + delta;
    }
    const Acc t = block_acc(a, lacc);
    const double n = t.n > 0.0 ? t.n : 1.0;
    const double mx = t.sx / n, my = t.sy / n, md = t.sd / n;
    // second pass over this thread's own rows (written above by the same thread)
    double cxx = 0.0, cyy = 0.0, cdd = 0.0, cxd = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBeamWG) {
        const double x = pps[i] - mx, y = pps[2 * ptot + i] - my, d = pps[5 * ptot + i] - md;
        cxx = cxx + x * x;
        cyy = cyy + y * y;
        cdd = cdd + d * d;
        cxd = cxd + x * d;
    }
    cxx = block_sum(cxx, lsum);
    cyy = block_sum(cyy, lsum);
    cdd = block_sum(cdd, lsum);
    cxd = block_sum(cxd, lsum);
    if (threadIdx.x == 0) {
        double* o = stats + 12 * g;
        o[0] = t.n;
        o[1] = t.mnx;
        o[2] = t.mxx;
        o[3] = t.mny;
        o[4] = t.mxy;
        o[5] = mx;
        o[6] = my;
        o[7] = md;
        o[8] = sqrt(cxx / n);
        o[9] = sqrt(cyy / n);
        o[10] = cdd;
        o[11] = cxd;
    }
}

// dJ/d(final state) of J_g = w0 rms_x + w1 rms_y + w2 |D| per group (D = cxd / cdd):
//   dJ/dx_i = w0 (x_i - mx) / (n rms_x) + w2 sign(D) (delta_i - md) / cdd,  dJ/dy_i = w1 (y_i - my) / (n rms_y),
//   dJ/dp_i = dJ/dx_i XI + dJ/dy_i YI,  dJ/dv_i = 0 (delta comes from the particle's momentum).
__global__ __launch_bounds__(kBeamWG) void k_objective_seed(const double* __restrict__ ideal, int64_t ndesign,
                                                            const double* __restrict__ pps,
                                                            const double* __restrict__ stats, int64_t ptot,
                                                            const int64_t* __restrict__ group_off,
                                                            const int32_t* __restrict__ group_design,
                                                            double w0, double w1, double w2,
                                                            double* __restrict__ adj) {
    const int64_t g = blockIdx.x;
    int64_t lo, hi;
    group_range(group_off, g, ptot, lo, hi);
    const double* id = ideal_of(ideal, group_design, g, ndesign);
    const double* XI = id + 7;
    const double* YI = id + 10;
    const double* s = stats + 12 * g;
    const double n = s[0] > 0.0 ? s[0] : 1.0;
    const double kx = s[8] > 0.0 ? w0 / (n * s[8]) : 0.0;
    const double ky = s[9] > 0.0 ? w1 / (n * s[9]) : 0.0;
    const double D = s[10] > 0.0 ? s[11] / s[10] : 0.0;
    const double kd = s[10] > 0.0 ? w2 * (D > 0.0 ? 1.0 : (D < 0.0 ? -1.0 : 0.0)) / s[10] : 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBeamWG) {
        const double gx = kx * (pps[i] - s[5]) + kd * (pps[5 * ptot + i] - s[7]);
        const double gy = ky * (pps[2 * ptot + i] - s[6]);
        for (int c = 0; c < 3; ++c) {
            adj[c * ptot + i] = gx * XI[c] + gy * YI[c];
            adj[(3 + c) * ptot + i] = 0.0;
        }
    }
}
#pragma clang fp contract(fast)

}  // namespace
}  // namespace cctamd_detail

using namespace cctamd_detail;

extern "C" {

int cctamd_phase_space(const double* ideal, int64_t ndesign, const double* state, const double* speed,
                       const double* momentum, int64_t nparticles, const int64_t* group_off,
                       const int32_t* group_design, int64_t ngroups, double* pps, double* stats, void* stream) {
    if (ndesign <= 0 || nparticles < 0 || ngroups < 0) return CCTAMD_EINVAL;
    if (ngroups == 0) return CCTAMD_OK;
    if (ngroups > 0x7fffffff) return CCTAMD_ERANGE;
    if (!ideal || !group_off || !group_design || !stats) return CCTAMD_EINVAL;
    if (nparticles > 0 && (!state || !speed || !momentum || !pps)) return CCTAMD_EINVAL;
    hipLaunchKernelGGL(k_phase_space, dim3((unsigned)ngroups), dim3(kBeamWG), 0, (hipStream_t)stream, ideal,
                       ndesign, state, speed, momentum, nparticles, group_off, group_design, pps, stats);
    return hip_status(hipGetLastError());
}

int cctamd_beam_objective_seed(const double* ideal, int64_t ndesign, const double* pps, const double* stats,
                               int64_t nparticles, const int64_t* group_off, const int32_t* group_design,
                               int64_t ngroups, const double* weights, double* adj, void* stream) {
    if (ndesign <= 0 || nparticles < 0 || ngroups < 0) return CCTAMD_EINVAL;
    if (ngroups == 0 || nparticles == 0) return CCTAMD_OK;
    if (ngroups > 0x7fffffff) return CCTAMD_ERANGE;
    if (!ideal || !pps || !stats || !group_off || !group_design || !weights || !adj) return CCTAMD_EINVAL;
    hipLaunchKernelGGL(k_objective_seed, dim3((unsigned)ngroups), dim3(kBeamWG), 0, (hipStream_t)stream, ideal,
                       ndesign, pps, stats, nparticles, group_off, group_design, weights[0], weights[1], weights[2],
                       adj);
    return hip_status(hipGetLastError());
}

}  // extern "C"
