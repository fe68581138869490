// cctamd: Biot-Savart field and RK4 tracking of protons through discretised
// CCT windings, written for MI355X (gfx950, CDNA4, wave64).
//
// Kernel map (one HIP stream, no allocation inside the library):
//   k_winding_generate  CCT parameters -> FP64 segments in the global frame
//   k_tile_origins      per 32-segment tile: FP64 origin (mean midpoint) and extent
//   k_pack_fp32         FP64 segments  -> FP32 tiles re-centred on the tile
//                                         origin (+ K/C rows)
//   k_bs_partial_*<N>   split-K Biot-Savart: one workgroup = one point block x
//                       one chunk of tiles; its waves (8 in the scalar-load
//                       kernels) share the block and interleave tiles; segments
//                       are wave-uniform (scalar loads into SGPRs), points live
//                       in VGPRs (N per lane)
//   k_bs_partial_f32k_row  32-point blocks: 16-lane rows walk their own tiles,
//                       segments broadcast within a row by DPP
//   k_field_reduce      fixed-order sum of the chunk partials + QS + uniform
//   k_rk4_stage<S>      same reduction, then RK4 stage S of each particle
//
// The adjoint kernels are in adjoint.hip; the C ABI is declared in
// include/cctamd.h.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>

#include "common.hpp"

namespace cctamd_detail {
namespace {

// ------------------------------------------------------------ geometry
#pragma clang fp contract(off)
__global__ void k_winding_generate(const double* __restrict__ params, const int64_t* __restrict__ count_off,
                                   const int64_t* __restrict__ dst, int64_t ncct, int64_t nseg_total,
                                   double* __restrict__ seg64) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nseg_total) return;
    int64_t lo = 0, hi = ncct - 1;  // last c with count_off[c] <= g
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (count_off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const double* prm = params + lo * CCTAMD_CCT_PARAMS;
    const int64_t i = g - count_off[lo];
    double p0[3], p1[3];
    cct_point(prm, i, p0);
    cct_point(prm, i + 1, p1);
    const double coef = prm[19];
    const double kl[3] = {coef * (p1[0] - p0[0]), coef * (p1[1] - p0[1]), coef * (p1[2] - p0[2])};
    const double m[3] = {0.5 * (p1[0] + p0[0]), 0.5 * (p1[1] + p0[1]), 0.5 * (p1[2] + p0[2])};
    const double* O = prm;
    const double* XI = prm + 3;
    const double* YI = prm + 6;
    const double* ZI = prm + 9;
    double* out = seg64 + (dst[lo] + i) * 8;
    for (int c = 0; c < 3; ++c) {
        out[c] = XI[c] * kl[0] + YI[c] * kl[1] + ZI[c] * kl[2];
        out[4 + c] = O[c] + (XI[c] * m[0] + YI[c] * m[1] + ZI[c] * m[2]);
    }
    out[3] = 0.0;
    out[7] = 0.0;
}
__global__ void k_winding_points(const double* __restrict__ params, int64_t npoints, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoints) return;
    double p[3];
    cct_point(params, i, p);
    out[3 * i] = p[0];
    out[3 * i + 1] = p[1];
    out[3 * i + 2] = p[2];
}

// Tile origin c = mean of the tile's current-carrying midpoints (a tile of pads
// keeps its first midpoint) and tile extent max |m - c|^2 over those rows.
__global__ void k_tile_origins(const double* __restrict__ seg64, int64_t ntile, double* __restrict__ tile_org) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntile) return;
    const double* rows = seg64 + t * kTile * 8;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int n = 0;
    for (int k = 0; k < kTile; ++k) {
        const double* r = rows + k * 8;
        if (r[0] == 0.0 && r[1] == 0.0 && r[2] == 0.0) continue;
        sx += r[4];
        sy += r[5];
        sz += r[6];
        ++n;
    }
    double cx = rows[4], cy = rows[5], cz = rows[6];
    if (n > 0) {
        cx = sx / n;
        cy = sy / n;
        cz = sz / n;
    }
    double ext = 0.0;
    for (int k = 0; k < kTile; ++k) {
        const double* r = rows + k * 8;
        if (r[0] == 0.0 && r[1] == 0.0 && r[2] == 0.0) continue;
        const double dx = r[4] - cx, dy = r[5] - cy, dz = r[6] - cz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        ext = d2 > ext ? d2 : ext;
    }
    tile_org[4 * t] = cx;
    tile_org[4 * t + 1] = cy;
    tile_org[4 * t + 2] = cz;
    tile_org[4 * t + 3] = ext;
}

__global__ void k_pack_fp32(const double* __restrict__ seg64, int64_t nseg, float* __restrict__ seg32,
                            const double* __restrict__ tile_org, float* __restrict__ seg32k) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    const int64_t t = i / kTile;
    const double* s = seg64 + i * 8;
    const double* o = tile_org + 4 * t;
    const double mx = s[4] - o[0], my = s[5] - o[1], mz = s[6] - o[2];
    float* d = seg32 + i * 8;
    d[0] = (float)s[0];
    d[1] = (float)s[1];
    d[2] = (float)s[2];
    d[3] = 0.0f;
    d[4] = (float)mx;
    d[5] = (float)my;
    d[6] = (float)mz;
    d[7] = 0.0f;
    if (seg32k) {
        // {kl, kl x (m - c), -2 (m - c), |m - c|^2}, each rounded once from FP64
        float* e = seg32k + i * 12;
        e[0] = (float)s[0];
        e[1] = (float)s[1];
        e[2] = (float)s[2];
        e[3] = (float)(s[1] * mz - s[2] * my);
        e[4] = (float)(s[2] * mx - s[0] * mz);
        e[5] = (float)(s[0] * my - s[1] * mx);
        e[6] = (float)(-2.0 * mx);
        e[7] = (float)(-2.0 * my);
        e[8] = (float)(-2.0 * mz);
        e[9] = (float)(mx * mx + my * my + mz * mz);
        e[10] = e[11] = 0.0f;
    }
}

#pragma clang fp contract(fast)

// ----------------------------------------------------- Biot-Savart core
// 1/sqrt(x) in FP64 for a normal x > 0: the hardware estimate y refined by one
// third-order step of the series (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + O(e^3),
// e = 1 - x y^2.  No special-value handling: a point on a segment midpoint
// (x = 0) gives a NaN term, as the reference's norm(r)**(-3) times a zero
// cross product does.
__device__ __forceinline__ double rsqrt_f64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// FP64 pair terms of tile t (seg64 rows) for NP points, added to FP64 accumulators:
// the term and summation order of k_bs_partial_f64.  The FP32 kernels take this
// form for a tile that one of their points lies among (kFp64Tile below): there the
// offset q = p - c from the tile origin is comparable with the distance to the
// nearest segment, and rounding q (or m - c) to FP32 moves that segment's term by
// ~1e-6 of itself, more than the whole FP32 sum may differ from the FP64 one.
// t is wave-uniform in the scalar-load kernels (scalar loads) and row- or
// group-uniform in the split-wave kernels (vector loads of a few addresses).
template <int NP>
__device__ __forceinline__ void tile_terms_f64(const double* __restrict__ seg64, int64_t t, const double (&px)[NP],
                                               const double (&py)[NP], const double (&pz)[NP], double (&ax)[NP],
                                               double (&ay)[NP], double (&az)[NP]) {
    const double2* s = reinterpret_cast<const double2*>(seg64) + t * kTile * 4;
    // one point at a time and no unrolling: this path is rare, and its registers count
    // against the occupancy of the FP32 loop of the same kernel
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        double bx = ax[j], by = ay[j], bz = az[j];
#pragma unroll 1
        for (int k = 0; k < kTile; ++k) {
            const double2 k01 = s[4 * k], k2_ = s[4 * k + 1], m01 = s[4 * k + 2], m2_ = s[4 * k + 3];
            const double rx = px[j] - m01.x, ry = py[j] - m01.y, rz = pz[j] - m2_.x;
            const double r2 = rx * rx + ry * ry + rz * rz;
            const double inv = rsqrt_f64(r2);
            const double inv3 = inv * inv * inv;
            bx += (k01.y * rz - k2_.x * ry) * inv3;
            by += (k2_.x * rx - k01.x * rz) * inv3;
            bz += (k01.x * ry - k01.y * rx) * inv3;
        }
        ax[j] = bx;
        ay[j] = by;
        az[j] = bz;
    }
}

// tile_terms_f64 for a lane group of LG lanes that share the tile (the split-wave
// kernels): in each pass every lane loads one segment's FP64 row (the loads of a
// pass go out together, and each row is read once per group), and the group's
// lanes take the segments in order from one another by lane shuffles.  gl is the
// lane's index in its group; the group's lanes are consecutive and all active.
template <int LG, int NP>
__device__ __forceinline__ void group_tile_terms_f64(const double* __restrict__ seg64, int64_t t, int gl,
                                                     const double (&px)[NP], const double (&py)[NP],
                                                     const double (&pz)[NP], double (&ax)[NP], double (&ay)[NP],
                                                     double (&az)[NP]) {
    const int first = (int)(threadIdx.x & (kWave - 1)) - gl;  // lane 0 of the group in the wave
    // one point at a time (the shuffles repeat per point): this path is rare, and its
    // registers count against the occupancy of the kernel's FP32 loop
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        double bx = ax[j], by = ay[j], bz = az[j];
#pragma unroll 1
        for (int k0 = 0; k0 < kTile; k0 += LG) {
            const double* r = seg64 + (t * kTile + k0 + gl) * 8;
            const double2 k01 = *reinterpret_cast<const double2*>(r), m01 = *reinterpret_cast<const double2*>(r + 4);
            const double k2 = r[2], m2 = r[6];
#pragma unroll 1
            for (int k = 0; k < LG; ++k) {
                const int src = first + k;
                const double kx = __shfl(k01.x, src), ky = __shfl(k01.y, src), kz = __shfl(k2, src);
                const double rx = px[j] - __shfl(m01.x, src), ry = py[j] - __shfl(m01.y, src);
                const double rz = pz[j] - __shfl(m2, src);
                const double r2 = rx * rx + ry * ry + rz * rz;
                const double inv = rsqrt_f64(r2);
                const double inv3 = inv * inv * inv;
                bx += (ky * rz - kz * ry) * inv3;
                by += (kz * rx - kx * rz) * inv3;
                bz += (kx * ry - ky * rx) * inv3;
            }
        }
        ax[j] = bx;
        ay[j] = by;
        az[j] = bz;
    }
}

// Tiles close to a point.  With e the tile extent (largest |m - c| of its
// segments, c the tile origin) and q = p - c:
//  * |q|^2 < 1.5 e^2 (kFp64Tile, |q| < 1.22 e): the point is among the tile's
//    segments; rounding q or m - c to FP32 moves the nearest segment's term by
//    ~1e-6 of itself, so the tile is summed in FP64 (tile_terms_f64);
//  * up to |q| < 2.5 e (kNearTile): the K x q - C form would cancel by up to
//    ~7x, so the K/C kernels sum the tile in the FP32 direct form (kl x r) |r|^-3;
//  * farther: the kernel's own form.
// Each decision is taken for the whole wave (or lane group) when any of its
// points qualifies.  A CPU emulation of the FP32 sums at the 300 worst of 129,024
// states recorded along the AG-CCT section's track gives the same largest field
// error (3.5e-7) for FP64 bands from |q|^2 < 1.25 e^2 up to 6.25 e^2, and 8.5e-7 at
// e^2; tiles of coarse, strongly tilted windings (360 segments per winding) reach
// 5 cm, which puts the in-bore beam of the single-CCT and population workloads at
// 1.4-1.6 e from the nearest tile origins, outside this band.
constexpr double kNearTile = 6.25;
constexpr double kFp64Tile = 1.5;

// Occupancy (waves per SIMD) of the scalar-load Biot-Savart kernels without the
// fused stage, which their fused variants (S >= 0) are held to: the stage
// update at the end alone would otherwise set their register count (~120-160
// VGPRs against 44 for k_bs_partial_f32k<1, 8>).  1 = no hint.
constexpr int kKindF32 = 0, kKindF32k = 1, kKindF64 = 2;
constexpr int fused_waves(int kind, int ppl, int s) {
    if (s < 0) return 1;
    if (kind == kKindF32) return ppl == 1 ? 6 : 4;
    if (kind == kKindF32k) return ppl == 1 ? 8 : ppl == 2 ? 6 : 4;
    return ppl == 1 ? 8 : ppl == 2 ? 5 : 4;
}

// ------------------------------------------------------- RK4 stage update
// Everything the stage update of a particle reads and writes.  scratch rows:
// 0-2 stage position (read by the Biot-Savart kernels), 3-5 stage velocity,
// 6-8 running sum of position slopes, 9-11 of velocity slopes, 12-13 the
// particle's partial-sum slot (so that the update loads it at once), 14 the
// arrival counters of the fused epilogue (int32, indexed by a point block's
// first point; zeroed by k_rk4_init, reset by the last arrival).
struct StageArgs {
    const double* partial;
    const double* qs;
    const int64_t* design_qs_off;
    const double* bconst;
    double* state;
    double* scratch;
    int64_t ptot;
    const double* kq;
    const double* dt;
    const int64_t* nsteps;
    int64_t step;
    double* record;
    int64_t record_every;
    double* tape;
};

__device__ inline int* arrival_counters(const StageArgs& sa) { return (int*)(sa.scratch + 14 * sa.ptot); }

#pragma clang fp contract(off)
// RK4 stage S of final_code/packages/base_utils.py:303-309 with the callback
// of particles.py:526-535 for particle g, by one whole wave (its lanes stride
// the chunk partials; lane 0 writes): a = (e/m) * (v x B), k = [v, a];
// Y1 = Y0 + dt/2 k1, Y2 = Y0 + dt/2 k2, Y3 = Y0 + dt k3,
// Y0 += dt/6 (((k1 + 2 k2) + 2 k3) + k4).
// LEAN (the fused epilogue): the chunk sum first, with two loads in flight per
// lane, and the particle's state loaded by lane 0 only after it, so that the
// stage update needs few registers beside the Biot-Savart loop's.  Otherwise
// every load is issued before the chunk sum (latency first, registers are
// plentiful in k_rk4_stage).  Both forms compute the same values.
template <int S, bool LEAN = false>
__device__ inline void rk4_stage_point(const StageArgs& sa, int64_t g) {
    const int64_t P = sa.ptot;
    double* ps = sa.scratch;
    double* vs = sa.scratch + 3 * P;
    double* kp = sa.scratch + 6 * P;
    double* kv = sa.scratch + 9 * P;
    double p[3], v[3], p0[3], v0[3], sp[3], sv[3], b[3];
    const PointSlot sl = load_slot(sa.scratch, P, g, sa.partial);
    int64_t ns;
    double k, h;
    AnalyticSet an;
    const auto load_state = [&]() {
        an = analytic_set(sl.design, sa.design_qs_off, sa.bconst);
        ns = sa.nsteps[g];
        k = sa.kq[g];
        h = sa.dt[g];
        for (int c = 0; c < 3; ++c) {
            const int64_t o = c * P + g;
            p[c] = ps[o];
            v[c] = vs[o];
            p0[c] = sa.state[o];
            v0[c] = sa.state[3 * P + o];
            if (S != 0) {
                sp[c] = kp[o];
                sv[c] = kv[o];
            }
        }
    };
    if (LEAN) {
        chunk_sum<2>(sl, b, sa.scratch);
        if ((threadIdx.x & (kWave - 1)) != 0) return;
        load_state();
        if (sa.step >= ns) return;
    } else {
        // The partial-sum slot first (the chunk loads depend on it), then every other load;
        // the finished-particle test comes after the chunk sum, so that the compiler does not
        // wait for nsteps before issuing the rest (finished particles keep their state).
        load_state();
        chunk_sum(sl, b, sa.scratch);
        if (sa.step >= ns) return;
        if ((threadIdx.x & (kWave - 1)) != 0) return;
    }
    analytic_fields(an, sa.qs, p, b);
    if (sa.tape) {  // stage point, velocity and field for the adjoint
        double* t = sa.tape + S * 9 * P;
        for (int c = 0; c < 3; ++c) {
            t[c * P + g] = p[c];
            t[(3 + c) * P + g] = v[c];
            t[(6 + c) * P + g] = b[c];
        }
    }
    const double cr[3] = {v[1] * b[2] - v[2] * b[1], -v[0] * b[2] + v[2] * b[0], v[0] * b[1] - v[1] * b[0]};
    const double a[3] = {cr[0] * k, cr[1] * k, cr[2] * k};
    const double h2 = h / 2;
    const bool rec = sa.record && ((sa.step + 1) % sa.record_every == 0);
    for (int c = 0; c < 3; ++c) {
        const int64_t o = c * P + g;
        if (S == 0) {
            kp[o] = v[c];
            kv[o] = a[c];
            ps[o] = p0[c] + v[c] * h2;
            vs[o] = v0[c] + a[c] * h2;
        } else if (S == 1) {
            kp[o] = sp[c] + v[c] * 2;
            kv[o] = sv[c] + a[c] * 2;
            ps[o] = p0[c] + v[c] * h2;
            vs[o] = v0[c] + a[c] * h2;
        } else if (S == 2) {
            kp[o] = sp[c] + v[c] * 2;
            kv[o] = sv[c] + a[c] * 2;
            ps[o] = p0[c] + v[c] * h;
            vs[o] = v0[c] + a[c] * h;
        } else {
            const double h6 = h / 6;
            const double np = p0[c] + (sp[c] + v[c]) * h6;
            const double nv = v0[c] + (sv[c] + a[c]) * h6;
            sa.state[o] = np;
            sa.state[3 * P + o] = nv;
            ps[o] = np;
            vs[o] = nv;
            if (rec) {
                double* r = sa.record + ((sa.step + 1) / sa.record_every) * 6 * P;
                r[o] = np;
                r[3 * P + o] = nv;
            }
        }
    }
}
#pragma clang fp contract(fast)

// Fused RK4 stage (S >= 0) at the end of a scalar-load Biot-Savart workgroup,
// after it has written its partials: the workgroup that completes a point
// block's set of chunks (the last of its w.nchunk items to arrive at the
// block's counter) sums the block's chunk partials and updates its particles,
// exactly as k_rk4_stage<S> would, so that a stage is one launch instead of
// two.  No workgroup waits for another: the others simply end.  Every wave
// waits until its partial stores are written to its XCD's L2, then after the
// barrier one agent-scope release by thread 0 writes them back before the
// count.  Only the last arrival acquires (its L2 and L1 lines invalidated),
// so that its loads see every other item's stores while the other workgroups
// keep their cached segment rows.
template <int S, int NW>
__device__ inline void stage_after_last_chunk(const Work& w, const StageArgs& sa) {
    __shared__ int last;
    __builtin_amdgcn_s_waitcnt(0);  // every counter, vector memory (the partial stores) included
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        int* ctr = arrival_counters(sa) + w.pt_start;
        last = atomicAdd(ctr, 1) == w.nchunk - 1;
        if (last) atomicExch(ctr, 0);  // ready for the next launch (all items of the block have arrived)
    }
    __syncthreads();
    if (!last) return;  // workgroup-uniform
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int i = threadIdx.x >> 6; i < w.pt_count; i += NW) rk4_stage_point<S, true>(sa, (int64_t)w.pt_start + i);
}

// FP32 terms on tiles re-centred at an FP64 origin; each tile's 32 terms are
// summed in FP32 and the tile sums are accumulated in FP64; near tiles in FP64.
template <int PPL, int S = -1>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(fused_waves(kKindF32, PPL, S)))) void k_bs_partial_f32(const float* __restrict__ seg32,
                                                         const double* __restrict__ tile_org,
                                                         const double* __restrict__ seg64,
                                                         const Work* __restrict__ work,
                                                         const double* __restrict__ pts, int64_t ptot,
                                                         double* __restrict__ partial, StageArgs sa) {
    constexpr int PB = kWave * PPL;
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double px[PPL], py[PPL], pz[PPL];
    load_points<PPL>(pts, ptot, w.pt_start, w.pt_count, lane, px, py, pz);
    double ax[PPL], ay[PPL], az[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) ax[j] = ay[j] = az[j] = 0.0;

    const int tend = w.tile_start + w.tile_count;
    for (int t = w.tile_start + wave; t < tend; t += kWavesPerWG) {
        const double ox = tile_org[4 * t], oy = tile_org[4 * t + 1], oz = tile_org[4 * t + 2];
        const double f64_2 = kFp64Tile * tile_org[4 * t + 3];
        float qx[PPL], qy[PPL], qz[PPL], sx[PPL], sy[PPL], sz[PPL];
        bool near = false;
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const double dx = px[j] - ox, dy = py[j] - oy, dz = pz[j] - oz;
            near = near || dx * dx + dy * dy + dz * dz < f64_2;
            qx[j] = (float)dx;
            qy[j] = (float)dy;
            qz[j] = (float)dz;
            sx[j] = sy[j] = sz[j] = 0.0f;
        }
        if (__any(near)) {  // wave-uniform
            tile_terms_f64<PPL>(seg64, t, px, py, pz, ax, ay, az);
            continue;
        }
        const float4* s = reinterpret_cast<const float4*>(seg32) + (int64_t)t * kTile * 2;
#pragma unroll 8
        for (int k = 0; k < kTile; ++k) {
            const float4 kl = s[2 * k];
            const float4 m = s[2 * k + 1];
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const float rx = qx[j] - m.x, ry = qy[j] - m.y, rz = qz[j] - m.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv3 = inv * inv * inv;
                sx[j] += (kl.y * rz - kl.z * ry) * inv3;
                sy[j] += (kl.z * rx - kl.x * rz) * inv3;
                sz[j] += (kl.x * ry - kl.y * rx) * inv3;
            }
        }
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            ax[j] += (double)sx[j];
            ay[j] += (double)sy[j];
            az[j] += (double)sz[j];
        }
    }
    write_partial<PPL>(ax, ay, az, wave, lane, partial + partial_base(w, PB), w.nchunk);
    if constexpr (S >= 0) stage_after_last_chunk<S, kWavesPerWG>(w, sa);
}

// FP32 tracking variant: kl x (q - m) = kl x q - kl x m, so within a tile the
// sums K = sum w kl and C = sum w (kl x m) (w = |q - m|^-3) give the tile field
// K x q - C; and |q - m|^2 = (|q|^2 + |m|^2) - 2 q.m needs no difference
// vector.  13 VALU slots per pair (add, 3 fma, rsq, 2 mul, 6 fma) instead of 17.
// |q|^2 is formed in FP64 per tile.  Both rewrites cancel when the point is
// close to the tile (|q| of the order of the tile extent e): for |q| >= 2.5 e
// the factors (|q|^2 + e^2) / |r|^2 and |q| / |r| stay below 3.3; nearer tiles
// are summed in the direct form or in FP64 (see kNearTile).

// Waves per workgroup of the scalar-load Biot-Savart kernels (K/C FP32 and FP64).
// Scalar loads return out of order, so a wave waits for all of its segment loads
// at once and cannot overlap them with the next ones; more resident waves per
// SIMD (with the same number of split-K chunks) cover those waits.  8 waves
// measured ~10 % faster than 4 on the 368,640-segment AG-CCT section, 16 no better.
#ifndef CCTAMD_BS_WAVES
#define CCTAMD_BS_WAVES 8
#endif
constexpr int kBsWaves = CCTAMD_BS_WAVES;

// Segments per iteration of the K/C, 16-point K/C, FP64 and split-wave FP64
// tile loops.  Measured on the AG-CCT section: 4 is 1.5 % faster than 8 for the
// K/C kernel (16: 1.5 % slower with one point per lane, 9 % with two); 2 or 8
// for the others stayed within 1 % of 4.
#ifndef CCTAMD_KC_UNROLL
#define CCTAMD_KC_UNROLL 4
#endif
#ifndef CCTAMD_KCG_UNROLL
#define CCTAMD_KCG_UNROLL 4
#endif
#ifndef CCTAMD_F64_UNROLL
#define CCTAMD_F64_UNROLL 4
#endif
#ifndef CCTAMD_F64G_UNROLL
#define CCTAMD_F64G_UNROLL 4
#endif

// One segment's K/C row as the K/C loop uses it (the row's last two floats are
// padding): a = kl.xyz, (kl x m).x; b = (kl x m).yz, -2m.xy; c = -2m.z, |m|^2.
struct KCSeg {
    float4 a, b;
    float2 c;
};

__device__ __forceinline__ KCSeg kc_seg(const float4* __restrict__ s, int k) {
    KCSeg g;
    g.a = s[3 * k];
    g.b = s[3 * k + 1];
    g.c = *reinterpret_cast<const float2*>(s + 3 * k + 2);
    return g;
}

// The K/C pair terms of one segment for a lane's PPL points (13 VALU slots each).
template <int PPL>
__device__ __forceinline__ void kc_terms(const KCSeg& g, const float (&qx)[PPL], const float (&qy)[PPL],
                                         const float (&qz)[PPL], const float (&qq)[PPL], float (&kx)[PPL],
                                         float (&ky)[PPL], float (&kz)[PPL], float (&cx)[PPL], float (&cy)[PPL],
                                         float (&cz)[PPL]) {
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        const float r2 = __builtin_fmaf(qz[j], g.c.x, __builtin_fmaf(qy[j], g.b.w, __builtin_fmaf(qx[j], g.b.z, qq[j] + g.c.y)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float w3 = inv * inv * inv;
        kx[j] += g.a.x * w3;
        ky[j] += g.a.y * w3;
        kz[j] += g.a.z * w3;
        cx[j] += g.a.w * w3;
        cy[j] += g.b.x * w3;
        cz[j] += g.b.y * w3;
    }
}

// One FP32 direct-form term (kl x r) |r|^-3 from a K/C row, r = q - m.
__device__ inline void tile_direct_term(const float4& a, const float4& b, const float4& c, float qx, float qy,
                                        float qz, float& bx, float& by, float& bz) {
    // m = -0.5 (-2m) is exact in FP32
    const float rx = __builtin_fmaf(0.5f, b.z, qx), ry = __builtin_fmaf(0.5f, b.w, qy), rz = __builtin_fmaf(0.5f, c.x, qz);
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float w3 = inv * inv * inv;
    bx += (a.y * rz - a.z * ry) * w3;
    by += (a.z * rx - a.x * rz) * w3;
    bz += (a.x * ry - a.y * rx) * w3;
}

template <int PPL, int NW, int S = -1>
__global__ __launch_bounds__(kWave * NW) __attribute__((amdgpu_waves_per_eu(fused_waves(kKindF32k, PPL, S)))) void k_bs_partial_f32k(const float* __restrict__ seg32k,
                                                          const double* __restrict__ tile_org,
                                                          const double* __restrict__ seg64,
                                                          const Work* __restrict__ work,
                                                          const double* __restrict__ pts, int64_t ptot,
                                                          double* __restrict__ partial, StageArgs sa) {
    constexpr int PB = kWave * PPL;
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double px[PPL], py[PPL], pz[PPL];
    load_points<PPL>(pts, ptot, w.pt_start, w.pt_count, lane, px, py, pz);
    double ax[PPL], ay[PPL], az[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) ax[j] = ay[j] = az[j] = 0.0;

    const int tend = w.tile_start + w.tile_count;
    for (int t = w.tile_start + wave; t < tend; t += NW) {
        const double ox = tile_org[4 * t], oy = tile_org[4 * t + 1], oz = tile_org[4 * t + 2];
        const double e2 = tile_org[4 * t + 3];
        float qx[PPL], qy[PPL], qz[PPL], qq[PPL], kx[PPL], ky[PPL], kz[PPL], cx[PPL], cy[PPL], cz[PPL];
        bool near = false, f64 = false;
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const double dx = px[j] - ox, dy = py[j] - oy, dz = pz[j] - oz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            near = near || d2 < kNearTile * e2;
            f64 = f64 || d2 < kFp64Tile * e2;
            qx[j] = (float)dx;
            qy[j] = (float)dy;
            qz[j] = (float)dz;
            qq[j] = (float)d2;
            kx[j] = ky[j] = kz[j] = cx[j] = cy[j] = cz[j] = 0.0f;
        }
        const float4* s = reinterpret_cast<const float4*>(seg32k) + (int64_t)t * kTile * 3;
        if (__any(f64)) {  // wave-uniform: FP64 terms for the whole wave on this tile
            tile_terms_f64<PPL>(seg64, t, px, py, pz, ax, ay, az);
            continue;
        }
        if (__any(near)) {  // wave-uniform: FP32 direct form
#pragma unroll 1
            for (int k = 0; k < kTile; ++k) {
                const float4 a = s[3 * k], b = s[3 * k + 1], c = s[3 * k + 2];
#pragma unroll
                for (int j = 0; j < PPL; ++j) tile_direct_term(a, b, c, qx[j], qy[j], qz[j], kx[j], ky[j], kz[j]);
            }
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                ax[j] += (double)kx[j];
                ay[j] += (double)ky[j];
                az[j] += (double)kz[j];
            }
            continue;
        }
#pragma unroll CCTAMD_KC_UNROLL
        for (int k = 0; k < kTile; ++k) kc_terms<PPL>(kc_seg(s, k), qx, qy, qz, qq, kx, ky, kz, cx, cy, cz);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            ax[j] += (double)((ky[j] * qz[j] - kz[j] * qy[j]) - cx[j]);
            ay[j] += (double)((kz[j] * qx[j] - kx[j] * qz[j]) - cy[j]);
            az[j] += (double)((kx[j] * qy[j] - ky[j] * qx[j]) - cz[j]);
        }
    }
    write_partial<PPL, NW>(ax, ay, az, wave, lane, partial + partial_base(w, PB), w.nchunk);
    if constexpr (S >= 0) stage_after_last_chunk<S, NW>(w, sa);
}

// 32-point blocks without LDS: each 16-lane row of a wave is a lane group with
// two points per lane that walks its own tiles.  Lane i of a row holds the K/C
// rows of segments i and i + 16 of the row's tile in registers (coalesced
// vector loads), and segment k reaches the whole row through the row-broadcast
// DPP control (lane k of each row to all 16 lanes) on the instruction that
// consumes the value, no LDS traffic and no barriers.  Registers bound its
// occupancy: with the tile loaded at the top of each round and a 6-wave hint it
// measured faster than prefetching the next tile (fewer waves per SIMD).
struct KCRow {
    float4 a, b, c;  // kl.xyz, C.x | C.yz, -2m.xy | -2m.z, |m|^2, 0, 0
};

// |q - m|^2 of segment S (lane S of the row) for both points of each lane, as
// ((|q|^2 + |m|^2) + qx (-2m.x)) + qy (-2m.y) + qz (-2m.z) like k_bs_partial_f32k,
// in one assembly block: the compiler pads every dependence between two separate
// inline-assembly statements with a wait state, which these plain accumulator
// dependences do not need (their broadcast sources are loaded, not computed).
template <int S>
__device__ __forceinline__ void r2_bcast(const KCRow& g, const float (&qx)[2], const float (&qy)[2],
                                         const float (&qz)[2], const float (&qq)[2], float& r0, float& r1) {
    asm volatile(
        "v_add_f32_dpp %0, %2, %3 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %2, %4 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %5, %6 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %5, %7 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %8, %9 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %8, %10 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %11, %12 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %11, %13 row_newbcast:%14 row_mask:0xf bank_mask:0xf"
        : "=&v"(r0), "=&v"(r1)
        : "v"(g.c.y), "v"(qq[0]), "v"(qq[1]), "v"(g.b.z), "v"(qx[0]), "v"(qx[1]), "v"(g.b.w), "v"(qy[0]),
          "v"(qy[1]), "v"(g.c.x), "v"(qz[0]), "v"(qz[1]), "i"(S));
}

template <int S>
__device__ __forceinline__ void kc_row_steps(const KCRow& g, const float (&qx)[2], const float (&qy)[2],
                                             const float (&qz)[2], const float (&qq)[2], float (&kx)[2],
                                             float (&ky)[2], float (&kz)[2], float (&cx)[2], float (&cy)[2],
                                             float (&cz)[2]) {
    float r0, r1;
    r2_bcast<S>(g, qx, qy, qz, qq, r0, r1);
    const float i0 = __builtin_amdgcn_rsqf(r0), i1 = __builtin_amdgcn_rsqf(r1);
    const float w0 = i0 * i0 * i0, w1 = i1 * i1 * i1;
    fmac_bcast<S>(kx[0], g.a.x, w0);
    fmac_bcast<S>(kx[1], g.a.x, w1);
    fmac_bcast<S>(ky[0], g.a.y, w0);
    fmac_bcast<S>(ky[1], g.a.y, w1);
    fmac_bcast<S>(kz[0], g.a.z, w0);
    fmac_bcast<S>(kz[1], g.a.z, w1);
    fmac_bcast<S>(cx[0], g.a.w, w0);
    fmac_bcast<S>(cx[1], g.a.w, w1);
    fmac_bcast<S>(cy[0], g.b.x, w0);
    fmac_bcast<S>(cy[1], g.b.x, w1);
    fmac_bcast<S>(cz[0], g.b.y, w0);
    fmac_bcast<S>(cz[1], g.b.y, w1);
    if constexpr (S + 1 < 16) kc_row_steps<S + 1>(g, qx, qy, qz, qq, kx, ky, kz, cx, cy, cz);
}

template <int S>
__device__ __forceinline__ void direct_row_steps(const KCRow& g, const float (&qx)[2], const float (&qy)[2],
                                                 const float (&qz)[2], float (&bx)[2], float (&by)[2],
                                                 float (&bz)[2]) {
    const KCRow h{make_float4(row_bcast<S>(g.a.x), row_bcast<S>(g.a.y), row_bcast<S>(g.a.z), 0.0f),
                  make_float4(0.0f, 0.0f, row_bcast<S>(g.b.z), row_bcast<S>(g.b.w)),
                  make_float4(row_bcast<S>(g.c.x), 0.0f, 0.0f, 0.0f)};
#pragma unroll
    for (int j = 0; j < 2; ++j) tile_direct_term(h.a, h.b, h.c, qx[j], qy[j], qz[j], bx[j], by[j], bz[j]);
    if constexpr (S + 1 < 16) direct_row_steps<S + 1>(g, qx, qy, qz, bx, by, bz);
}

// 6 waves per SIMD (80 VGPRs): hints of 7 (72 VGPRs, 10 spilled) and 8 (64, 23
// spilled) measured 8 % and 11 % slower on the 64-design population.
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(6))) void k_bs_partial_f32k_row(const float* __restrict__ seg32k,
                                                             const double* __restrict__ tile_org,
                                                             const double* __restrict__ seg64,
                                                             const Work* __restrict__ work,
                                                             const double* __restrict__ pts, int64_t ptot,
                                                             double* __restrict__ partial) {
    constexpr int LG = 16;                  // lanes per row
    constexpr int PB = 2 * LG;              // 32 points per block
    constexpr int NS = kWavesPerWG * 4;     // tile streams (rows) per workgroup
    __shared__ double red[kWavesPerWG][3][PB];
    // each lane's FP64 sums live in LDS (its own slots, no synchronisation): updated once
    // per tile, they would otherwise hold 12 registers through the FP32 loop
    __shared__ double acc[6][kWG];  // [component * 2 + point][thread]
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int gl = lane & (LG - 1);
    const int stream = wave * 4 + (lane >> 4);
    const int tid = threadIdx.x;
    double px[2], py[2], pz[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        int idx = gl + LG * j;
        idx = idx < w.pt_count ? idx : 0;  // surplus lanes duplicate point 0, results discarded
        const int64_t gp = (int64_t)w.pt_start + idx;
        px[j] = pts[gp];
        py[j] = pts[ptot + gp];
        pz[j] = pts[2 * ptot + gp];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c][tid] = 0.0;
    const int tend = w.tile_start + w.tile_count;
    const float4* src = reinterpret_cast<const float4*>(seg32k);
    auto load = [&](int t, KCRow& g0, KCRow& g1) {  // segments gl and gl + 16 of tile t
        const float4* r0 = src + ((int64_t)t * kTile + gl) * 3;
        const float4* r1 = r0 + LG * 3;
        g0 = KCRow{r0[0], r0[1], r0[2]};
        g1 = KCRow{r1[0], r1[1], r1[2]};
    };
    int t = w.tile_start + stream;
    while (t < tend) {  // uniform within a row
        const int tn = t + NS;
        KCRow g0, g1;
        load(t, g0, g1);
        const double ox = tile_org[4 * t], oy = tile_org[4 * t + 1], oz = tile_org[4 * t + 2];
        const double e2 = tile_org[4 * t + 3];
        float qx[2], qy[2], qz[2], qq[2], kx[2], ky[2], kz[2], cx[2], cy[2], cz[2];
        bool near = false, f64 = false;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double dx = px[j] - ox, dy = py[j] - oy, dz = pz[j] - oz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            near = near || d2 < kNearTile * e2;
            f64 = f64 || d2 < kFp64Tile * e2;
            qx[j] = (float)dx;
            qy[j] = (float)dy;
            qz[j] = (float)dz;
            qq[j] = (float)d2;
            kx[j] = ky[j] = kz[j] = cx[j] = cy[j] = cz[j] = 0.0f;
        }
        if (__any(f64)) {  // FP64 terms for the active rows on this round, each row its own tile
            double ax[2] = {acc[0][tid], acc[1][tid]}, ay[2] = {acc[2][tid], acc[3][tid]};
            double az[2] = {acc[4][tid], acc[5][tid]};
            group_tile_terms_f64<LG, 2>(seg64, t, gl, px, py, pz, ax, ay, az);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[j][tid] = ax[j];
                acc[2 + j][tid] = ay[j];
                acc[4 + j][tid] = az[j];
            }
        } else if (__any(near)) {  // FP32 direct form for the active rows on this round
            direct_row_steps<0>(g0, qx, qy, qz, kx, ky, kz);
            direct_row_steps<0>(g1, qx, qy, qz, kx, ky, kz);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[j][tid] += (double)kx[j];
                acc[2 + j][tid] += (double)ky[j];
                acc[4 + j][tid] += (double)kz[j];
            }
        } else {
            kc_row_steps<0>(g0, qx, qy, qz, qq, kx, ky, kz, cx, cy, cz);
            kc_row_steps<0>(g1, qx, qy, qz, qq, kx, ky, kz, cx, cy, cz);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[j][tid] += (double)((ky[j] * qz[j] - kz[j] * qy[j]) - cx[j]);
                acc[2 + j][tid] += (double)((kz[j] * qx[j] - kx[j] * qz[j]) - cy[j]);
                acc[4 + j][tid] += (double)((kx[j] * qy[j] - ky[j] * qx[j]) - cz[j]);
            }
        }
        t = tn;
    }
    double ax[2] = {acc[0][tid], acc[1][tid]}, ay[2] = {acc[2][tid], acc[3][tid]}, az[2] = {acc[4][tid], acc[5][tid]};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int off = LG; off < kWave; off <<= 1) {  // add the rows (fixed order)
            ax[j] += __shfl_xor(ax[j], off);
            ay[j] += __shfl_xor(ay[j], off);
            az[j] += __shfl_xor(az[j], off);
        }
    }
    if (lane < LG) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            red[wave][0][lane + LG * j] = ax[j];
            red[wave][1][lane + LG * j] = ay[j];
            red[wave][2][lane + LG * j] = az[j];
        }
    }
    __syncthreads();
    double* out = partial + partial_base(w, PB);
    for (int k = threadIdx.x; k < 3 * PB; k += kWG) {
        const int c = k / PB, i = k % PB;
        out[(int64_t)(i * 3 + c) * w.nchunk] = ((red[0][c][i] + red[1][c][i]) + red[2][c][i]) + red[3][c][i];
    }
}

// Small point blocks (32 or 16 points): each wave is split into G lane groups
// that hold the same points (PL per lane) and walk different tiles, so every
// lane works.  Each wave reads 3 ds_read_b128 per segment whatever PL is, and
// four waves share the CU's LDS, so two points per lane (4 groups for 32-point
// blocks, 8 for 16-point blocks) halve the LDS reads per pair of one point per
// lane; four points per lane cost more in registers (occupancy) than they save.
// The workgroup stages 4G tiles of K/C rows in LDS per round; a lane reads
// the segments of its group's tile with LDS broadcasts (G addresses per wave),
// which issue outside the VALU.  Missing tiles of the last round are
// zero-current segments far away.
template <int G, int PL>
__global__ __launch_bounds__(kWG) void k_bs_partial_f32k_grp(const float* __restrict__ seg32k,
                                                              const double* __restrict__ tile_org,
                                                              const double* __restrict__ seg64,
                                                              const Work* __restrict__ work,
                                                              const double* __restrict__ pts, int64_t ptot,
                                                              double* __restrict__ partial) {
    constexpr int LG = kWave / G;         // lanes per group
    constexpr int PB = LG * PL;           // points per block: PL per lane
    constexpr int NS = kWavesPerWG * G;   // tile streams per workgroup
    constexpr int ROW = kTile * 3;        // float4 per tile
    __shared__ float4 lds[NS][ROW + 1];   // +1: the streams start in different banks
    __shared__ double red[kWavesPerWG][3][PB];
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int gl = lane % LG;
    const int stream = wave * G + lane / LG;
    double px[PL], py[PL], pz[PL], ax[PL], ay[PL], az[PL];
#pragma unroll
    for (int j = 0; j < PL; ++j) {
        int idx = gl + LG * j;
        idx = idx < w.pt_count ? idx : 0;  // surplus lanes duplicate point 0, results discarded
        const int64_t gp = (int64_t)w.pt_start + idx;
        px[j] = pts[gp];
        py[j] = pts[ptot + gp];
        pz[j] = pts[2 * ptot + gp];
        ax[j] = ay[j] = az[j] = 0.0;
    }
    const int tend = w.tile_start + w.tile_count;
    const float4* src = reinterpret_cast<const float4*>(seg32k);
    // missing tiles: zero current, |m|^2 = 1e6 keeps |q - m|^2 > 0
    const auto pad = [](int e) {
        return e % 3 == 2 ? make_float4(0.0f, 1.0e6f, 0.0f, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    for (int t0 = w.tile_start; t0 < tend; t0 += NS) {
        stage_tiles<NS, ROW>(lds, src, t0, tend, pad);
        __syncthreads();
        const bool real = t0 + stream < tend;
        const int t = real ? t0 + stream : w.tile_start;
        const double ox = tile_org[4 * t], oy = tile_org[4 * t + 1], oz = tile_org[4 * t + 2];
        const double e2 = tile_org[4 * t + 3];
        float qx[PL], qy[PL], qz[PL], qq[PL], kx[PL], ky[PL], kz[PL], cx[PL], cy[PL], cz[PL];
        bool near = false, f64 = false;
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            const double dx = px[j] - ox, dy = py[j] - oy, dz = pz[j] - oz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            near = near || (real && d2 < kNearTile * e2);
            f64 = f64 || (real && d2 < kFp64Tile * e2);
            qx[j] = (float)dx;
            qy[j] = (float)dy;
            qz[j] = (float)dz;
            qq[j] = (float)d2;
            kx[j] = ky[j] = kz[j] = cx[j] = cy[j] = cz[j] = 0.0f;
        }
        const float4* r4 = lds[stream];
        if (__any(f64)) {  // FP64 terms for this round, each lane group its own tile (if any)
            if (real) group_tile_terms_f64<LG, PL>(seg64, t, gl, px, py, pz, ax, ay, az);
            __syncthreads();
            continue;
        }
        if (__any(near)) {  // FP32 direct form for this round
#pragma unroll 1
            for (int k = 0; k < kTile; ++k) {
                const float4 a = r4[3 * k], b = r4[3 * k + 1], c = r4[3 * k + 2];
#pragma unroll
                for (int j = 0; j < PL; ++j) tile_direct_term(a, b, c, qx[j], qy[j], qz[j], kx[j], ky[j], kz[j]);
            }
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                ax[j] += (double)kx[j];
                ay[j] += (double)ky[j];
                az[j] += (double)kz[j];
            }
            __syncthreads();
            continue;
        }
#pragma unroll CCTAMD_KCG_UNROLL
        for (int k = 0; k < kTile; ++k) {
            const float4 a = r4[3 * k];
            const float4 b = r4[3 * k + 1];
            const float4 c = r4[3 * k + 2];
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                const float r2 = __builtin_fmaf(qz[j], c.x, __builtin_fmaf(qy[j], b.w, __builtin_fmaf(qx[j], b.z, qq[j] + c.y)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float w3 = inv * inv * inv;
                kx[j] += a.x * w3;
                ky[j] += a.y * w3;
                kz[j] += a.z * w3;
                cx[j] += a.w * w3;
                cy[j] += b.x * w3;
                cz[j] += b.y * w3;
            }
        }
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            ax[j] += (double)((ky[j] * qz[j] - kz[j] * qy[j]) - cx[j]);
            ay[j] += (double)((kz[j] * qx[j] - kx[j] * qz[j]) - cy[j]);
            az[j] += (double)((kx[j] * qy[j] - ky[j] * qx[j]) - cz[j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < PL; ++j) {
#pragma unroll
        for (int off = LG; off < kWave; off <<= 1) {  // add the lane groups (fixed order)
            ax[j] += __shfl_xor(ax[j], off);
            ay[j] += __shfl_xor(ay[j], off);
            az[j] += __shfl_xor(az[j], off);
        }
    }
    if (lane < LG) {
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            red[wave][0][lane + LG * j] = ax[j];
            red[wave][1][lane + LG * j] = ay[j];
            red[wave][2][lane + LG * j] = az[j];
        }
    }
    __syncthreads();
    double* out = partial + partial_base(w, PB);
    for (int k = threadIdx.x; k < 3 * PB; k += kWG) {
        const int c = k / PB, i = k % PB;
        out[(int64_t)(i * 3 + c) * w.nchunk] = ((red[0][c][i] + red[1][c][i]) + red[2][c][i]) + red[3][c][i];
    }
}

// Plain FP64 terms on global-frame segments.
template <int PPL, int NW, int S = -1>
__global__ __launch_bounds__(kWave * NW) __attribute__((amdgpu_waves_per_eu(fused_waves(kKindF64, PPL, S)))) void k_bs_partial_f64(const double* __restrict__ seg64,
                                                         const Work* __restrict__ work,
                                                         const double* __restrict__ pts, int64_t ptot,
                                                         double* __restrict__ partial, StageArgs sa) {
    constexpr int PB = kWave * PPL;
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double px[PPL], py[PPL], pz[PPL];
    load_points<PPL>(pts, ptot, w.pt_start, w.pt_count, lane, px, py, pz);
    double ax[PPL], ay[PPL], az[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) ax[j] = ay[j] = az[j] = 0.0;

    const int tend = w.tile_start + w.tile_count;
    for (int t = w.tile_start + wave; t < tend; t += NW) {
        const double2* s = reinterpret_cast<const double2*>(seg64) + (int64_t)t * kTile * 4;
#pragma unroll CCTAMD_F64_UNROLL
        for (int k = 0; k < kTile; ++k) {
            const double2 k01 = s[4 * k], k2_ = s[4 * k + 1], m01 = s[4 * k + 2], m2_ = s[4 * k + 3];
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const double rx = px[j] - m01.x, ry = py[j] - m01.y, rz = pz[j] - m2_.x;
                const double r2 = rx * rx + ry * ry + rz * rz;
                const double inv = rsqrt_f64(r2);
                const double inv3 = inv * inv * inv;
                ax[j] += (k01.y * rz - k2_.x * ry) * inv3;
                ay[j] += (k2_.x * rx - k01.x * rz) * inv3;
                az[j] += (k01.x * ry - k01.y * rx) * inv3;
            }
        }
    }
    write_partial<PPL, NW>(ax, ay, az, wave, lane, partial + partial_base(w, PB), w.nchunk);
    if constexpr (S >= 0) stage_after_last_chunk<S, NW>(w, sa);
}

// FP64 split-wave form of k_bs_partial_f64 for designs with few particles
// (points-per-lane codes -2 / -4): G lane groups of 64 / G lanes hold the same
// PL points per lane and walk different tiles, FP64 segment rows staged in LDS
// and read as broadcasts, so no lane idles where a 64-point block would be
// half empty.  Same pair term and summation order per point as k_bs_partial_f64.
// Tiles are staged in PARTS parts (32 / PARTS segments each), which keeps the
// LDS per workgroup small enough for several workgroups per CU (<4, 2, 2>:
// 7 waves per SIMD, 2.5 % faster than whole tiles at 4; 8 groups of 4 points
// per lane were slower).
template <int G, int PL, int PARTS>
__global__ __launch_bounds__(kWG) void k_bs_partial_f64_grp(const double* __restrict__ seg64,
                                                             const Work* __restrict__ work,
                                                             const double* __restrict__ pts, int64_t ptot,
                                                             double* __restrict__ partial) {
    constexpr int LG = kWave / G;         // lanes per group
    constexpr int PB = LG * PL;           // points per block
    constexpr int NS = kWavesPerWG * G;   // tile streams per workgroup
    constexpr int ROW = kTile * 4;        // float4 (one double2 each) per tile in seg64
    constexpr int SEGS = kTile / PARTS;   // segments per staged part of a tile
    // LDS rows drop seg64's two pad doubles: (kl.x, kl.y), (kl.z, m.x), (m.y, m.z)
    __shared__ float4 lds[NS][3 * SEGS + 1];
    __shared__ double red[kWavesPerWG][3][PB];
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int gl = lane % LG;
    const int stream = wave * G + lane / LG;
    double px[PL], py[PL], pz[PL], ax[PL], ay[PL], az[PL];
#pragma unroll
    for (int j = 0; j < PL; ++j) {
        int idx = gl + LG * j;
        idx = idx < w.pt_count ? idx : 0;  // surplus lanes duplicate point 0, results discarded
        const int64_t gp = (int64_t)w.pt_start + idx;
        px[j] = pts[gp];
        py[j] = pts[ptot + gp];
        pz[j] = pts[2 * ptot + gp];
        ax[j] = ay[j] = az[j] = 0.0;
    }
    const int tend = w.tile_start + w.tile_count;
    const float4* src = reinterpret_cast<const float4*>(seg64);
    static_assert(NS * SEGS == kWG, "one segment per thread and staging round");
    const int st = (int)threadIdx.x / SEGS, sk = (int)threadIdx.x % SEGS;
    for (int t0 = w.tile_start; t0 < tend; t0 += NS) {
#pragma unroll 1
        for (int h = 0; h < PARTS; ++h) {
            float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0, v2 = v0, v3 = v0;  // absent tiles: not read
            if (t0 + st < tend) {
                const float4* g = src + (int64_t)(t0 + st) * ROW + (h * SEGS + sk) * 4;
                v0 = g[0];  // kl.x, kl.y
                v1 = g[1];  // kl.z, pad
                v2 = g[2];  // m.x, m.y
                v3 = g[3];  // m.z, pad
            }
            lds[st][3 * sk] = v0;
            lds[st][3 * sk + 1] = make_float4(v1.x, v1.y, v2.x, v2.y);
            lds[st][3 * sk + 2] = make_float4(v2.z, v2.w, v3.x, v3.y);
            __syncthreads();
            if (t0 + stream < tend) {  // uniform within a lane group
                const double2* r = reinterpret_cast<const double2*>(lds[stream]);
#pragma unroll CCTAMD_F64G_UNROLL
                for (int k = 0; k < SEGS; ++k) {
                    const double2 k01 = r[3 * k], k2m0 = r[3 * k + 1], m12 = r[3 * k + 2];
#pragma unroll
                    for (int j = 0; j < PL; ++j) {
                        const double rx = px[j] - k2m0.y, ry = py[j] - m12.x, rz = pz[j] - m12.y;
                        const double r2 = rx * rx + ry * ry + rz * rz;
                        const double inv = rsqrt_f64(r2);
                        const double inv3 = inv * inv * inv;
                        ax[j] += (k01.y * rz - k2m0.x * ry) * inv3;
                        ay[j] += (k2m0.x * rx - k01.x * rz) * inv3;
                        az[j] += (k01.x * ry - k01.y * rx) * inv3;
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < PL; ++j) {
#pragma unroll
        for (int off = LG; off < kWave; off <<= 1) {  // add the lane groups (fixed order)
            ax[j] += __shfl_xor(ax[j], off);
            ay[j] += __shfl_xor(ay[j], off);
            az[j] += __shfl_xor(az[j], off);
        }
    }
    if (lane < LG) {
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            red[wave][0][lane + LG * j] = ax[j];
            red[wave][1][lane + LG * j] = ay[j];
            red[wave][2][lane + LG * j] = az[j];
        }
    }
    __syncthreads();
    double* out = partial + partial_base(w, PB);
    for (int k = threadIdx.x; k < 3 * PB; k += kWG) {
        const int c = k / PB, i = k % PB;
        out[(int64_t)(i * 3 + c) * w.nchunk] = ((red[0][c][i] + red[1][c][i]) + red[2][c][i]) + red[3][c][i];
    }
}

// ------------------------------------------------ reduction + QS + RK4
__global__ __launch_bounds__(kWG) void k_field_reduce(int ppl, const double* __restrict__ partial,
                                                      const int64_t* __restrict__ design_wbase,
                                                      const int32_t* __restrict__ design_nchunk,
                                                      const int64_t* __restrict__ design_pt_off, int64_t nd,
                                                      const double* __restrict__ qs,
                                                      const int64_t* __restrict__ design_qs_off,
                                                      const double* __restrict__ bconst,
                                                      const double* __restrict__ pts, int64_t ptot,
                                                      double* __restrict__ B) {
    const int64_t g = (int64_t)blockIdx.x * kWavesPerWG + (threadIdx.x >> 6);
    if (g >= ptot) return;
    const double p[3] = {pts[g], pts[ptot + g], pts[2 * ptot + g]};
    const PointSlot sl = point_slot(g, ppl, partial, design_wbase, design_nchunk, design_pt_off, nd);
    double b[3];
    chunk_sum(sl, b, pts);
    if ((threadIdx.x & (kWave - 1)) != 0) return;
    analytic_fields(sl.design, qs, design_qs_off, bconst, p, b);
    B[g] = b[0];
    B[ptot + g] = b[1];
    B[2 * ptot + g] = b[2];
}

// Stage state of every particle from its state (scratch rows as in StageArgs).
__global__ void k_rk4_init(const double* __restrict__ state, int64_t ptot, double* __restrict__ scratch,
                           double* __restrict__ record, int ppl, const double* __restrict__ partial,
                           const int64_t* __restrict__ design_wbase, const int32_t* __restrict__ design_nchunk,
                           const int64_t* __restrict__ design_pt_off, int64_t nd) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ptot) return;
    for (int c = 0; c < 6; ++c) {
        scratch[c * ptot + g] = state[c * ptot + g];
        if (record) record[c * ptot + g] = state[c * ptot + g];
    }
    store_slot(scratch, ptot, g, point_slot(g, ppl, partial, design_wbase, design_nchunk, design_pt_off, nd),
               partial);
    scratch[14 * ptot + g] = 0.0;  // arrival counters of the fused stages (two int32 per double)
}

// RK4 stage S of every particle, one wave per particle (the stage update itself
// is rk4_stage_point, shared with the fused Biot-Savart epilogue).
template <int S>
__global__ __launch_bounds__(kWG) void k_rk4_stage(StageArgs sa) {
    const int64_t g = (int64_t)blockIdx.x * kWavesPerWG + (threadIdx.x >> 6);
    if (g >= sa.ptot) return;
    rk4_stage_point<S>(sa, g);
}

template <int PPL, int S = -1>
hipError_t launch_bs(int precision, const void* seg, const double* tile_org, const double* seg_near, const Work* work,
                     int64_t nwork, const double* pts, int64_t ptot, double* partial, hipStream_t st,
                     const StageArgs& sa = StageArgs{}) {
    if (nwork == 0) return hipSuccess;
    if (precision == 32)
        hipLaunchKernelGGL((k_bs_partial_f32<PPL, S>), dim3((unsigned)nwork), dim3(kWG), 0, st,
                           (const float*)seg, tile_org, seg_near, work, pts, ptot, partial, sa);
    else if (precision == 33)
        hipLaunchKernelGGL((k_bs_partial_f32k<PPL, kBsWaves, S>), dim3((unsigned)nwork), dim3(kWave * kBsWaves), 0,
                           st, (const float*)seg, tile_org, seg_near, work, pts, ptot, partial, sa);
    else
        hipLaunchKernelGGL((k_bs_partial_f64<PPL, kBsWaves, S>), dim3((unsigned)nwork), dim3(kWave * kBsWaves), 0,
                           st, (const double*)seg, work, pts, ptot, partial, sa);
    return hipGetLastError();
}

// One Biot-Savart pass at the stage points with RK4 stage S fused into the
// kernel's end (scalar-load kernels: points-per-lane codes 1, 2, 4).
template <int S>
hipError_t launch_bs_stage(const DeviceSet& ds, int64_t ptot, double* partial, const StageArgs& sa, hipStream_t st) {
    const double* pts = sa.scratch;  // rows 0-2: the stage points
    switch (ds.ppl) {
        case 1:
            return launch_bs<1, S>(ds.precision, ds.seg, ds.tile_org, ds.seg_near, ds.work, ds.nwork, pts, ptot,
                                   partial, st, sa);
        case 2:
            return launch_bs<2, S>(ds.precision, ds.seg, ds.tile_org, ds.seg_near, ds.work, ds.nwork, pts, ptot,
                                   partial, st, sa);
        default:
            return launch_bs<4, S>(ds.precision, ds.seg, ds.tile_org, ds.seg_near, ds.work, ds.nwork, pts, ptot,
                                   partial, st, sa);
    }
}

}  // namespace

// ------------------------------------------------------- shared launchers
bool valid_common(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                  int64_t nwork, int64_t ndesign) {
    if (precision != 32 && precision != 33 && precision != 64) return false;
    if (!valid_ppl(ppl) || (ppl < 0 && precision == 32)) return false;
    if (nwork < 0 || ndesign <= 0 || nwork > 0x7fffffff) return false;
    if (nwork > 0 && seg == nullptr) return false;
    if (nwork > 0 && precision != 64 && (tile_org == nullptr || seg_near == nullptr)) return false;
    return true;
}

hipError_t launch_bs_any(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                         const Work* work, int64_t nwork, const double* pts, int64_t ptot, double* partial,
                         hipStream_t st) {
    if (ppl < 0) {
        if (nwork == 0) return hipSuccess;
        if (precision == 64) {
            if (ppl == -2)
                hipLaunchKernelGGL((k_bs_partial_f64_grp<4, 2, 2>), dim3((unsigned)nwork), dim3(kWG), 0, st,
                                   (const double*)seg, work, pts, ptot, partial);
            else
                hipLaunchKernelGGL((k_bs_partial_f64_grp<8, 2, 4>), dim3((unsigned)nwork), dim3(kWG), 0, st,
                                   (const double*)seg, work, pts, ptot, partial);
            return hipGetLastError();
        }
        if (precision != 33) return hipErrorInvalidValue;
        if (ppl == -2)
            hipLaunchKernelGGL(k_bs_partial_f32k_row, dim3((unsigned)nwork), dim3(kWG), 0, st,
                               (const float*)seg, tile_org, seg_near, work, pts, ptot, partial);
        else
            hipLaunchKernelGGL((k_bs_partial_f32k_grp<8, 2>), dim3((unsigned)nwork), dim3(kWG), 0, st,
                               (const float*)seg, tile_org, seg_near, work, pts, ptot, partial);
        return hipGetLastError();
    }
    switch (ppl) {
        case 1: return launch_bs<1>(precision, seg, tile_org, seg_near, work, nwork, pts, ptot, partial, st);
        case 2: return launch_bs<2>(precision, seg, tile_org, seg_near, work, nwork, pts, ptot, partial, st);
        default: return launch_bs<4>(precision, seg, tile_org, seg_near, work, nwork, pts, ptot, partial, st);
    }
}

int field_checked(const DeviceSet& ds, const double* points, int64_t npoints, double* partial, double* B,
                  hipStream_t st) {
    if (!valid_common(ds.precision, ds.ppl, ds.seg, ds.tile_org, ds.seg_near, ds.nwork, ds.ndesign))
        return CCTAMD_EINVAL;
    if (npoints < 0) return CCTAMD_EINVAL;
    if (npoints == 0) return CCTAMD_OK;
    if (!points || !B || !ds.design_wbase || !ds.design_nchunk || !ds.design_pt_off || !ds.design_qs_off ||
        !ds.bconst)
        return CCTAMD_EINVAL;
    if (ds.nwork > 0 && (!ds.work || !partial)) return CCTAMD_EINVAL;
    const int64_t blocks = ceil_div(npoints, kWavesPerWG);
    if (blocks > 0x7fffffff) return CCTAMD_ERANGE;
    hipError_t e = launch_bs_any(ds.precision, ds.ppl, ds.seg, ds.tile_org, ds.seg_near, ds.work, ds.nwork, points,
                                 npoints, partial, st);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_field_reduce, dim3((unsigned)blocks), dim3(kWG), 0, st, ds.ppl, partial, ds.design_wbase,
                       ds.design_nchunk, ds.design_pt_off, ds.ndesign, ds.qs, ds.design_qs_off, ds.bconst, points,
                       npoints, B);
    return hip_status(hipGetLastError());
}

// Each RK4 stage is a Biot-Savart launch plus k_rk4_stage.  With
// CCTAMD_FUSED_STAGES=1 in the environment the scalar-load kernels
// (points-per-lane codes 1, 2, 4) run the stage at their end instead (one
// launch per stage, the same bits).  Not the default, being slower: the stage
// update runs at the end of the launch in one workgroup per point block, whose
// waves sum their points' chunk partials one point after another (the chunk
// order of k_rk4_stage), a serial tail of ~45 us per launch on the 23,040-segment
// single CCT whatever the item count (6x slower there; 2.9x on the
// 368,640-segment AG-CCT section), against ~5 us for the stage kernel.
bool fused_stages(const DeviceSet& ds) {
    if (ds.ppl <= 0 || ds.nwork == 0) return false;
    const char* env = getenv("CCTAMD_FUSED_STAGES");
    return env && env[0] == '1';
}

hipError_t launch_rk4_steps(const DeviceSet& ds, double* state, int64_t ptot, const double* kq, const double* dt,
                            const int64_t* nsteps, int64_t step0, int64_t step1, double* scratch, double* partial,
                            double* record, int64_t record_every, double* tape, hipStream_t st) {
    const int64_t eblocks = ceil_div(ptot, 256);
    const int64_t rblocks = ceil_div(ptot, kWavesPerWG);
    hipLaunchKernelGGL(k_rk4_init, dim3((unsigned)eblocks), dim3(256), 0, st, state, ptot, scratch,
                       step0 == 0 ? record : nullptr, ds.ppl, partial, ds.design_wbase, ds.design_nchunk,
                       ds.design_pt_off, ds.ndesign);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool fused = fused_stages(ds);
    StageArgs sa{partial, ds.qs, ds.design_qs_off, ds.bconst, state, scratch, ptot, kq, dt, nsteps, 0, record,
                 record_every, nullptr};
    for (int64_t step = step0; step < step1; ++step) {
        sa.step = step;
        sa.tape = tape ? tape + (step - step0) * 36 * ptot : nullptr;
#define CCTAMD_STAGE(S)                                                                                           \
    if (fused) {                                                                                                  \
        e = launch_bs_stage<S>(ds, ptot, partial, sa, st);                                                        \
    } else {                                                                                                      \
        e = launch_bs_any(ds.precision, ds.ppl, ds.seg, ds.tile_org, ds.seg_near, ds.work, ds.nwork, scratch,    \
                          ptot, partial, st);                                                                     \
        if (e != hipSuccess) return e;                                                                            \
        hipLaunchKernelGGL(k_rk4_stage<S>, dim3((unsigned)rblocks), dim3(kWG), 0, st, sa);                        \
        e = hipGetLastError();                                                                                    \
    }                                                                                                             \
    if (e != hipSuccess) return e;
        CCTAMD_STAGE(0)
        CCTAMD_STAGE(1)
        CCTAMD_STAGE(2)
        CCTAMD_STAGE(3)
#undef CCTAMD_STAGE
    }
    return hipSuccess;
}

}  // namespace cctamd_detail

using namespace cctamd_detail;

extern "C" {
int cctamd_version(void) { return 200; }

const char* cctamd_error_string(int code) {
    if (code == CCTAMD_OK) return "ok";
    if (code == CCTAMD_EINVAL) return "cctamd: invalid argument";
    if (code == CCTAMD_ERANGE) return "cctamd: size out of range";
    if (code > 0) return hipGetErrorString((hipError_t)code);
    return "cctamd: unknown error";
}

int cctamd_winding_generate(const double* params, const int64_t* count_off, const int64_t* dst, int64_t ncct,
                            int64_t nseg_total, double* seg64, void* stream) {
    if (ncct < 0 || nseg_total < 0) return CCTAMD_EINVAL;
    if (ncct == 0 || nseg_total == 0) return CCTAMD_OK;
    if (!params || !count_off || !dst || !seg64) return CCTAMD_EINVAL;
    const int threads = 256;
    const int64_t blocks = ceil_div(nseg_total, threads);
    if (blocks > 0x7fffffff) return CCTAMD_ERANGE;
    hipLaunchKernelGGL(k_winding_generate, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, params,
                       count_off, dst, ncct, nseg_total, seg64);
    return hip_status(hipGetLastError());
}

int cctamd_winding_points(const double* params, int64_t npoints, double* out, void* stream) {
    if (npoints < 0) return CCTAMD_EINVAL;
    if (npoints == 0) return CCTAMD_OK;
    if (!params || !out) return CCTAMD_EINVAL;
    const int64_t blocks = ceil_div(npoints, 256);
    if (blocks > 0x7fffffff) return CCTAMD_ERANGE;
    hipLaunchKernelGGL(k_winding_points, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, params, npoints,
                       out);
    return hip_status(hipGetLastError());
}

int cctamd_pack_fp32(const double* seg64, int64_t nseg, float* seg32, double* tile_org, float* seg32k,
                     void* stream) {
    if (nseg < 0 || nseg % kTile) return CCTAMD_EINVAL;
    if (nseg == 0) return CCTAMD_OK;
    if (!seg64 || !seg32 || !tile_org) return CCTAMD_EINVAL;
    const int threads = 256;
    const int64_t blocks = ceil_div(nseg, threads);
    if (blocks > 0x7fffffff) return CCTAMD_ERANGE;
    const int64_t ntile = nseg / kTile;
    hipLaunchKernelGGL(k_tile_origins, dim3((unsigned)ceil_div(ntile, threads)), dim3(threads), 0,
                       (hipStream_t)stream, seg64, ntile, tile_org);
    hipLaunchKernelGGL(k_pack_fp32, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, seg64, nseg,
                       seg32, tile_org, seg32k);
    return hip_status(hipGetLastError());
}

int64_t cctamd_plan(int64_t ndesign, const int64_t* design_ntile, const int64_t* design_npt,
                    const int64_t* design_tile_off, const int64_t* design_pt_off, int ppl, int64_t target_items,
                    int64_t min_tiles, int32_t* work, int64_t* design_wbase, int32_t* design_nchunk) {
    if (ndesign <= 0 || !design_ntile || !design_npt) return CCTAMD_EINVAL;
    if (!valid_ppl(ppl)) return CCTAMD_EINVAL;
    if (target_items <= 0 || min_tiles <= 0) return CCTAMD_EINVAL;
    const int64_t pb = pb_of(ppl);
    int64_t total = 0;
    for (int64_t d = 0; d < ndesign; ++d) {
        if (design_ntile[d] < 0 || design_npt[d] < 0) return CCTAMD_EINVAL;
        total += ceil_div(design_npt[d], pb) * design_ntile[d];
    }
    // Split-wave kernels: tiles per item a multiple of the workgroup's 16 tile streams, so that
    // every row gets as many tiles (2 % on the 64-design population).  Not for the scalar-load
    // kernels: there the item count should divide evenly over the CUs (2,048 items of 45 tiles
    // measured 4 % faster than 1,920 of 48 on the AG-CCT section).
    const int64_t streams = ppl > 0 ? 1 : 16;
    const int64_t per_item = ceil_div(std::max(min_tiles, ceil_div(std::max<int64_t>(total, 1), target_items)),
                                      streams) * streams;
    int64_t n = 0;
    for (int64_t d = 0; d < ndesign; ++d) {
        const int64_t nb = ceil_div(design_npt[d], pb);
        const int64_t T = design_ntile[d];
        // At least one item per point block, also for a design without tiles (an item of no
        // tiles writes zero partials): the fused RK4 stages update a block's particles at the
        // end of its items.
        int64_t nc = nb > 0 ? std::max<int64_t>(1, ceil_div(T, per_item)) : 0;
        // Workgroups go to the XCDs round-robin and item (b, c) is launched as
        // n + b nc + c: with nc a multiple of the XCD count the items of one tile
        // chunk (same segments, other point blocks) share an XCD and its L2.
        if (nb > 1 && ceil_div(nc, kXcds) * kXcds <= T) nc = ceil_div(nc, kXcds) * kXcds;
        if (work) {
            if (!design_tile_off || !design_pt_off || !design_wbase || !design_nchunk) return CCTAMD_EINVAL;
            if (nc > 0x7fffffff) return CCTAMD_ERANGE;
            design_wbase[d] = n;
            design_nchunk[d] = (int32_t)nc;
            for (int64_t b = 0; b < nb && nc > 0; ++b) {
                for (int64_t c = 0; c < nc; ++c) {
                    const int64_t t0 = T * c / nc, t1 = T * (c + 1) / nc;
                    const int64_t ps = design_pt_off[d] + b * pb;
                    const int64_t pc = std::min(pb, design_npt[d] - b * pb);
                    const int64_t ts = design_tile_off[d] + t0;
                    if (ps + pc > 0x7fffffff || ts + (t1 - t0) > 0x7fffffff) return CCTAMD_ERANGE;
                    int32_t* w = work + CCTAMD_WORK_INTS * (n + b * nc + c);
                    w[0] = (int32_t)d;
                    w[1] = (int32_t)ps;
                    w[2] = (int32_t)pc;
                    w[3] = (int32_t)ts;
                    w[4] = (int32_t)(t1 - t0);
                    w[5] = (int32_t)c;
                    w[6] = (int32_t)nc;
                    w[7] = (int32_t)b;
                }
            }
        }
        n += nb * nc;
    }
    return n;
}

int cctamd_bs_partial(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                      const int32_t* work, int64_t nwork, const double* points, int64_t npoints, double* partial,
                      void* stream) {
    if (!valid_common(precision, ppl, seg, tile_org, seg_near, nwork, 1)) return CCTAMD_EINVAL;
    if (npoints < 0) return CCTAMD_EINVAL;
    if (nwork == 0 || npoints == 0) return CCTAMD_OK;
    if (!work || !points || !partial) return CCTAMD_EINVAL;
    return hip_status(launch_bs_any(precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, points, npoints,
                                    partial, (hipStream_t)stream));
}

int cctamd_bs_field(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                    const int32_t* work, int64_t nwork, const int64_t* design_wbase, const int32_t* design_nchunk,
                    const int64_t* design_pt_off, int64_t ndesign, const double* qs, const int64_t* design_qs_off,
                    const double* bconst, const double* points, int64_t npoints, double* partial, double* B,
                    void* stream) {
    if (nwork > 0 && !partial) return CCTAMD_EINVAL;
    const DeviceSet ds{precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, design_wbase,
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, bconst};
    return field_checked(ds, points, npoints, partial, B, (hipStream_t)stream);
}

int cctamd_track_rk4(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                     const int32_t* work, int64_t nwork, const int64_t* design_wbase, const int32_t* design_nchunk,
                     const int64_t* design_pt_off, int64_t ndesign, const double* qs, const int64_t* design_qs_off,
                     const double* bconst, double* state, int64_t nparticles, const double* kq, const double* dt,
                     const int64_t* nsteps, int64_t max_steps, double* scratch, double* partial, double* record,
                     int64_t record_every, void* stream) {
    if (!valid_common(precision, ppl, seg, tile_org, seg_near, nwork, ndesign)) return CCTAMD_EINVAL;
    if (nparticles < 0 || max_steps < 0 || record_every <= 0) return CCTAMD_EINVAL;
    if (nparticles == 0) return CCTAMD_OK;
    if (!state || !kq || !dt || !nsteps || !scratch || !design_wbase || !design_nchunk || !design_pt_off ||
        !design_qs_off || !bconst)
        return CCTAMD_EINVAL;
    if (nwork > 0 && (!work || !partial)) return CCTAMD_EINVAL;
    if (ceil_div(nparticles, kWavesPerWG) > 0x7fffffff) return CCTAMD_ERANGE;
    const DeviceSet ds{precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, design_wbase,
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, bconst};
    return hip_status(launch_rk4_steps(ds, state, nparticles, kq, dt, nsteps, 0, max_steps, scratch, partial, record,
                                       record_every, nullptr, (hipStream_t)stream));
}

}  // extern "C"
