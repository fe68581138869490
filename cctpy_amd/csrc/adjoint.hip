// Adjoint (reverse-mode) of the RK4 tracking path for MI355X (gfx950).
//
// Discrete adjoint of one RK4 step Y' = Y + h/6 (k1 + 2 k2 + 2 k3 + k4) with
// k = [V, A], A = kq V x B(P) (final_code/packages/base_utils.py:303-309,
// particles.py:526-535), stages reversed 4 -> 1:
//   A_k-bar -> V_k-bar += kq (B_k x A_k-bar),  B_k-bar = kq (A_k-bar x V_k),
//   P_k-bar = J_B(P_k)^T B_k-bar  (backward Biot-Savart kernel + QS),
// and the parameter gradients of the field:
//   current of a CCT:   sum over its segments of B-bar . dB_i  (B is linear in I)
//   tilt coefficient j: sum_i dJ/dkl_i . dkl_i/dc_j + dJ/dm_i . dm_i/dc_j with
//                       dP/dc_j = dP/dphi * sin((j+1) ksi)  (cct.py:166-188, 213-224)
//   QS gradients:       analytic (magnets.py:637-668).
// The forward is recomputed from checkpoints every K steps; the recompute
// records a tape of stage points, velocities and fields.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace cctamd_detail {
namespace {

constexpr int kGradTilts = 4;          // tilt slots carried by the adjoint
constexpr int kTanRows = kTile + 1;    // winding points per tile in the tangent table
constexpr int kTileGrad = 1 + kGradTilts;

// ---------------------------------------------------------- tangents
// dP/dphi at ksi in the global frame ((-t sinh(eta) sin(phi), t sinh(eta) cos(phi), 0)
// locally) and sin((j+1) ksi), j = 0..3.
__device__ inline void cct_tangent(const double* __restrict__ prm, int64_t i, float out[8]) {
    const double ksi = prm[15] + prm[16] * (double)i;
    const double phi = cct_phi(prm, ksi);
    const double t = prm[12] / (prm[14] - cos(ksi));
    const double lx = -t * prm[13] * sin(phi), ly = t * prm[13] * cos(phi);
    const double* XI = prm + 3;
    const double* YI = prm + 6;
    for (int c = 0; c < 3; ++c) out[c] = (float)(XI[c] * lx + YI[c] * ly);
    const int ntilt = (int)prm[20];
    for (int j = 0; j < kGradTilts; ++j) out[3 + j] = j < ntilt ? (float)sin((double)(j + 1) * ksi) : 0.0f;
    out[7] = 0.0f;
}

__device__ inline void put_tangent(float* __restrict__ seg_tan, int64_t tile, int entry, const float v[8],
                                   float real) {
    float* o = seg_tan + (tile * kTanRows + entry) * 8;
    for (int c = 0; c < 7; ++c) o[c] = v[c];
    o[7] = real;
}

// One thread per winding point (count + 1 per CCT; every CCT starts on a tile
// boundary). Point i of a CCT whose segments start at row dst is the start of
// segment i (entry (dst+i) % kTile of its tile, real = 1) and, when segment i-1
// ends a tile, that tile's entry kTile; the end point of the last segment also
// fills the first pad slot of a partial tile (real = 0).
__global__ void k_winding_tangents(const double* __restrict__ params, const int64_t* __restrict__ count_off,
                                   const int64_t* __restrict__ dst, int64_t ncct, int64_t npoint_total,
                                   float* __restrict__ seg_tan) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npoint_total) return;
    int64_t lo = 0, hi = ncct - 1;  // last c with count_off[c] + c <= g
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (count_off[mid] + mid <= g) lo = mid; else hi = mid - 1;
    }
    const double* prm = params + lo * CCTAMD_CCT_PARAMS;
    const int64_t i = g - count_off[lo] - lo;
    const int64_t n = count_off[lo + 1] - count_off[lo];
    float v[8];
    cct_tangent(prm, i, v);
    const int64_t u = dst[lo] + i;
    if (i < n || u % kTile != 0) put_tangent(seg_tan, u / kTile, (int)(u % kTile), v, i < n ? 1.0f : 0.0f);
    if (i >= 1 && u % kTile == 0) put_tangent(seg_tan, u / kTile - 1, kTile, v, 0.0f);
}

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// -------------------------------------------------- backward Biot-Savart
// For each point p with weight b (the adjoint of the field at p) and segment i,
// r = p - m_i, w = |r|^-3:
//   point side:   pbar(p) += u = w (b x kl_i) - 3 w |r|^-2 (kl_i . (r x b)) r
//   segment side: dJ/dkl_i += w (r x b),  dJ/dm_i -= u,  dJ/dI += w kl_i . (r x b) / I
// The segment sums are contracted at once with the winding tangents: with
// kl_i = coef (P_i+1 - P_i), m_i = (P_i + P_i+1)/2 and dP_i/dc_j = T_i s_ij,
//   dJ/dc_j = sum over winding points i of s_ij T_i . (A_i-1 - B_i),
//   A = coef dJ/dkl + dJ/dm / 2,  B = coef dJ/dkl - dJ/dm / 2,
// so each point's end-of-segment and start-of-segment shares take one dot product.
// tgrad[block][tile] = {sum b . dB (current), dJ/dc_0..3 (tilt coefficients)}.
// FP32 terms on re-centred tiles, FP64 across tiles.

// Segments per iteration of the backward kernels' tile loops, as measured
// fastest: 64-point blocks 4 with one point per lane and 2 with more (the
// adjoint pass of a 256-proton AG-CCT section: 17 % faster than 2 with one
// point per lane, 3 % slower with two), 16-point blocks 2, 32-point blocks 1
// (5 % faster than 2 on the 64-design population).
#ifndef CCTAMD_BWD1_UNROLL
#define CCTAMD_BWD1_UNROLL 4
#endif
constexpr int bwd_unroll(int ppl) { return ppl == 1 ? CCTAMD_BWD1_UNROLL : 2; }
constexpr int kBwdGrpUnroll = 2;
constexpr int kBwdGrp2Unroll = 1;

// One point x segment term. hb = b / 2; returns c = r x b, w = |r|^-3,
// h = u / 2 and e = w kl . c.
struct BwdTerm {
    float cx, cy, cz, w, hx, hy, hz, e;
};

__device__ __forceinline__ BwdTerm bwd_term(float qx, float qy, float qz, float bx, float by, float bz, float hbx,
                                            float hby, float hbz, const float4 kl, const float4 m) {
    BwdTerm o;
    const float rx = qx - m.x, ry = qy - m.y, rz = qz - m.z;
    const float r2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv2 = inv * inv;
    o.w = inv2 * inv;
    o.cx = __builtin_fmaf(ry, bz, -(rz * by));
    o.cy = __builtin_fmaf(rz, bx, -(rx * bz));
    o.cz = __builtin_fmaf(rx, by, -(ry * bx));
    const float sdot = __builtin_fmaf(kl.z, o.cz, __builtin_fmaf(kl.y, o.cy, kl.x * o.cx));
    o.e = sdot * o.w;
    const float na = -1.5f * (o.e * inv2);
    const float kx = __builtin_fmaf(hby, kl.z, -(hbz * kl.y));
    const float ky = __builtin_fmaf(hbz, kl.x, -(hbx * kl.z));
    const float kz = __builtin_fmaf(hbx, kl.y, -(hby * kl.x));
    o.hx = __builtin_fmaf(na, rx, o.w * kx);
    o.hy = __builtin_fmaf(na, ry, o.w * ky);
    o.hz = __builtin_fmaf(na, rz, o.w * kz);
    return o;
}

// Segment-side step of the tangent contraction for one segment: gk = dJ/dkl and
// gh = -dJ/dm / 2 summed over this lane's points, ck = coef of the segment (0 for a
// pad slot), ta/tb = tangent entry of its start point. Adds the start point's share
// (A_prev - B) . T s_j to g and leaves A of this segment in a.
struct TanAcc {
    float ax, ay, az, g0, g1, g2, g3;
};

__device__ __forceinline__ void tan_step(TanAcc& t, float ck, float gkx, float gky, float gkz, float ghx, float ghy,
                                         float ghz, const float4 ta, const float4 tb) {
    const float dx = t.ax - __builtin_fmaf(ck, gkx, ghx);
    const float dy = t.ay - __builtin_fmaf(ck, gky, ghy);
    const float dz = t.az - __builtin_fmaf(ck, gkz, ghz);
    const float d = __builtin_fmaf(dz, ta.z, __builtin_fmaf(dy, ta.y, dx * ta.x));
    t.g0 = __builtin_fmaf(ta.w, d, t.g0);
    t.g1 = __builtin_fmaf(tb.x, d, t.g1);
    t.g2 = __builtin_fmaf(tb.y, d, t.g2);
    t.g3 = __builtin_fmaf(tb.z, d, t.g3);
    t.ax = __builtin_fmaf(ck, gkx, -ghx);
    t.ay = __builtin_fmaf(ck, gky, -ghy);
    t.az = __builtin_fmaf(ck, gkz, -ghz);
}

// End point of the tile's last segment (entry kTile).
__device__ __forceinline__ void tan_end(TanAcc& t, const float4 ta, const float4 tb) {
    const float d = __builtin_fmaf(t.az, ta.z, __builtin_fmaf(t.ay, ta.y, t.ax * ta.x));
    t.g0 = __builtin_fmaf(ta.w, d, t.g0);
    t.g1 = __builtin_fmaf(tb.x, d, t.g1);
    t.g2 = __builtin_fmaf(tb.y, d, t.g2);
    t.g3 = __builtin_fmaf(tb.z, d, t.g3);
}

template <int PPL>
__global__ __launch_bounds__(kWG) void k_bs_backward_f32(
    const float* __restrict__ seg32, const double* __restrict__ tile_org, const float* __restrict__ seg_tan,
    const int32_t* __restrict__ tile_cct, const double* __restrict__ cct_coef, const Work* __restrict__ work,
    const double* __restrict__ pts, const double* __restrict__ wts, int64_t ptot, int64_t ntile_total,
    double* __restrict__ partial, double* __restrict__ tgrad) {
    constexpr int PB = kWave * PPL;
    constexpr int kUnroll = bwd_unroll(PPL);
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double px[PPL], py[PPL], pz[PPL], wx[PPL], wy[PPL], wz[PPL];
    load_points<PPL>(pts, ptot, w.pt_start, w.pt_count, lane, px, py, pz);
    load_points<PPL>(wts, ptot, w.pt_start, w.pt_count, lane, wx, wy, wz);
#pragma unroll
    for (int j = 0; j < PPL; ++j) {  // surplus lanes carry no weight
        if (lane + kWave * j >= w.pt_count) wx[j] = wy[j] = wz[j] = 0.0;
    }
    float bx[PPL], by[PPL], bz[PPL], hbx[PPL], hby[PPL], hbz[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        bx[j] = (float)wx[j];
        by[j] = (float)wy[j];
        bz[j] = (float)wz[j];
        hbx[j] = 0.5f * bx[j];
        hby[j] = 0.5f * by[j];
        hbz[j] = 0.5f * bz[j];
    }
    double ax[PPL], ay[PPL], az[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) ax[j] = ay[j] = az[j] = 0.0;

    const int tend = w.tile_start + w.tile_count;
    for (int t = w.tile_start + wave; t < tend; t += kWavesPerWG) {
        const double ox = tile_org[4 * t], oy = tile_org[4 * t + 1], oz = tile_org[4 * t + 2];
        const int cct = tile_cct[t];
        const float coef = cct >= 0 ? (float)cct_coef[cct] : 0.0f;
        float qx[PPL], qy[PPL], qz[PPL], sx[PPL], sy[PPL], sz[PPL];
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            qx[j] = (float)(px[j] - ox);
            qy[j] = (float)(py[j] - oy);
            qz[j] = (float)(pz[j] - oz);
            sx[j] = sy[j] = sz[j] = 0.0f;
        }
        float cur = 0.0f;
        TanAcc ta{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const float4* s = reinterpret_cast<const float4*>(seg32) + (int64_t)t * kTile * 2;
        const float4* tn = reinterpret_cast<const float4*>(seg_tan) + (int64_t)t * kTanRows * 2;
#pragma unroll kUnroll
        for (int k = 0; k < kTile; ++k) {
            const float4 kl = s[2 * k];
            const float4 m = s[2 * k + 1];
            float gkx, gky, gkz, ghx, ghy, ghz;
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const BwdTerm b = bwd_term(qx[j], qy[j], qz[j], bx[j], by[j], bz[j], hbx[j], hby[j], hbz[j], kl, m);
                sx[j] += b.hx;
                sy[j] += b.hy;
                sz[j] += b.hz;
                cur += b.e;
                if (j == 0) {
                    gkx = b.w * b.cx;
                    gky = b.w * b.cy;
                    gkz = b.w * b.cz;
                    ghx = b.hx;
                    ghy = b.hy;
                    ghz = b.hz;
                } else {
                    gkx = __builtin_fmaf(b.w, b.cx, gkx);
                    gky = __builtin_fmaf(b.w, b.cy, gky);
                    gkz = __builtin_fmaf(b.w, b.cz, gkz);
                    ghx += b.hx;
                    ghy += b.hy;
                    ghz += b.hz;
                }
            }
            const float4 t0 = tn[2 * k], t1 = tn[2 * k + 1];
            tan_step(ta, coef * t1.w, gkx, gky, gkz, ghx, ghy, ghz, t0, t1);
        }
        tan_end(ta, tn[2 * kTile], tn[2 * kTile + 1]);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {  // s holds u / 2
            ax[j] += 2.0 * (double)sx[j];
            ay[j] += 2.0 * (double)sy[j];
            az[j] += 2.0 * (double)sz[j];
        }
        cur = wave_sum(cur);
        const float g0 = wave_sum(ta.g0), g1 = wave_sum(ta.g1), g2 = wave_sum(ta.g2), g3 = wave_sum(ta.g3);
        if (lane == 0 && cct >= 0) {
            double* tg = tgrad + ((int64_t)w.block * ntile_total + t) * kTileGrad;
            tg[0] += (double)cur;
            tg[1] += (double)g0;
            tg[2] += (double)g1;
            tg[3] += (double)g2;
            tg[4] += (double)g3;
        }
    }
    write_partial<PPL>(ax, ay, az, wave, lane, partial + partial_base(w, PB), w.nchunk);
}

// Split-wave form of k_bs_backward_f32 for 32- or 16-point blocks: each wave
// splits into G lane groups that hold the same points and walk different tiles
// (seg32 rows and tangents staged in LDS, read as broadcasts); the per-segment
// sums reduce within a group, and the point adjoints across groups at the end.
// It serves 16-point blocks (G = 4); 32-point blocks use k_bs_backward_f32_grp2<4>,
// 3 % faster on the 64-design population than this kernel with G = 2 (which ran
// 10 % faster with a 5-wave occupancy hint than without; the same hint made
// k_bs_backward_f32<2> 9 % slower, so that one has none).
template <int G>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(G == 2 ? 5 : 4))) void k_bs_backward_f32_grp(
    const float* __restrict__ seg32, const double* __restrict__ tile_org, const float* __restrict__ seg_tan,
    const int32_t* __restrict__ tile_cct, const double* __restrict__ cct_coef, const Work* __restrict__ work,
    const double* __restrict__ pts, const double* __restrict__ wts, int64_t ptot, int64_t ntile_total,
    double* __restrict__ partial, double* __restrict__ tgrad) {
    constexpr int PB = kWave / G;
    constexpr int NS = kWavesPerWG * G;  // tile streams per workgroup
    constexpr int SROW = kTile * 2;      // float4 per tile of seg32
    constexpr int TROW = kTanRows * 2;   // float4 per tile of tangents
    __shared__ float4 lseg[NS][SROW + 1];
    __shared__ float4 ltan[NS][TROW + 1];
    __shared__ double red[kWavesPerWG][3][PB];
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int gl = lane % PB;
    const int stream = wave * G + lane / PB;
    const bool has = gl < w.pt_count;
    const int64_t gp = (int64_t)w.pt_start + (has ? gl : 0);
    const double px = pts[gp], py = pts[ptot + gp], pz = pts[2 * ptot + gp];
    const float bx = has ? (float)wts[gp] : 0.0f;  // surplus lanes carry no weight
    const float by = has ? (float)wts[ptot + gp] : 0.0f;
    const float bz = has ? (float)wts[2 * ptot + gp] : 0.0f;
    const float hbx = 0.5f * bx, hby = 0.5f * by, hbz = 0.5f * bz;
    double ax = 0.0, ay = 0.0, az = 0.0;
    const int tend = w.tile_start + w.tile_count;
    const float4* s4 = reinterpret_cast<const float4*>(seg32);
    const float4* t4 = reinterpret_cast<const float4*>(seg_tan);
    for (int t0 = w.tile_start; t0 < tend; t0 += NS) {
        const auto zero = [](int) { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); };
        stage_tiles<NS, SROW>(lseg, s4, t0, tend, zero);
        stage_tiles<NS, TROW>(ltan, t4, t0, tend, zero);
        __syncthreads();
        const int t = t0 + stream;
        if (t < tend) {  // uniform within a lane group
            const int cct = tile_cct[t];
            const float coef = cct >= 0 ? (float)cct_coef[cct] : 0.0f;
            const float qx = (float)(px - tile_org[4 * t]);
            const float qy = (float)(py - tile_org[4 * t + 1]);
            const float qz = (float)(pz - tile_org[4 * t + 2]);
            float sx = 0.0f, sy = 0.0f, sz = 0.0f, cur = 0.0f;
            TanAcc ta{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            const float4* sr = lseg[stream];
            const float4* tr = ltan[stream];
#pragma unroll kBwdGrpUnroll
            for (int k = 0; k < kTile; ++k) {
                const BwdTerm b = bwd_term(qx, qy, qz, bx, by, bz, hbx, hby, hbz, sr[2 * k], sr[2 * k + 1]);
                sx += b.hx;
                sy += b.hy;
                sz += b.hz;
                cur += b.e;
                const float4 e0 = tr[2 * k], e1 = tr[2 * k + 1];
                tan_step(ta, coef * e1.w, b.w * b.cx, b.w * b.cy, b.w * b.cz, b.hx, b.hy, b.hz, e0, e1);
            }
            tan_end(ta, tr[2 * kTile], tr[2 * kTile + 1]);
            ax += 2.0 * (double)sx;  // s holds u / 2
            ay += 2.0 * (double)sy;
            az += 2.0 * (double)sz;
            float g0 = ta.g0, g1 = ta.g1, g2 = ta.g2, g3 = ta.g3;
#pragma unroll
            for (int off = PB / 2; off > 0; off >>= 1) {  // sum over the group's points
                cur += __shfl_xor(cur, off);
                g0 += __shfl_xor(g0, off);
                g1 += __shfl_xor(g1, off);
                g2 += __shfl_xor(g2, off);
                g3 += __shfl_xor(g3, off);
            }
            if (gl == 0 && cct >= 0) {
                double* tg = tgrad + ((int64_t)w.block * ntile_total + t) * kTileGrad;
                tg[0] += (double)cur;
                tg[1] += (double)g0;
                tg[2] += (double)g1;
                tg[3] += (double)g2;
                tg[4] += (double)g3;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int off = PB; off < kWave; off <<= 1) {  // add the lane groups (fixed order)
        ax += __shfl_xor(ax, off);
        ay += __shfl_xor(ay, off);
        az += __shfl_xor(az, off);
    }
    if (lane < PB) {
        red[wave][0][lane] = ax;
        red[wave][1][lane] = ay;
        red[wave][2][lane] = az;
    }
    __syncthreads();
    double* out = partial + partial_base(w, PB);
    for (int k = threadIdx.x; k < 3 * PB; k += kWG) {
        const int c = k / PB, i = k % PB;
        out[(int64_t)(i * 3 + c) * w.nchunk] = ((red[0][c][i] + red[1][c][i]) + red[2][c][i]) + red[3][c][i];
    }
}

// Split-wave form with two points per lane: G lane groups of 64 / G lanes, each
// lane holding points gl and gl + 64 / G of the block (32-point blocks for G = 4),
// so that the tangent contraction of a segment and its LDS reads are shared by
// two points (~50 VALU instructions per pair against ~58).  Its 16 tile streams
// take ~37 KB of LDS per workgroup, which holds it at 4 waves per SIMD.
template <int G>
__global__ __launch_bounds__(kWG) void k_bs_backward_f32_grp2(
    const float* __restrict__ seg32, const double* __restrict__ tile_org, const float* __restrict__ seg_tan,
    const int32_t* __restrict__ tile_cct, const double* __restrict__ cct_coef, const Work* __restrict__ work,
    const double* __restrict__ pts, const double* __restrict__ wts, int64_t ptot, int64_t ntile_total,
    double* __restrict__ partial, double* __restrict__ tgrad) {
    constexpr int LG = kWave / G;        // lanes per group
    constexpr int PB = 2 * LG;           // points per block
    constexpr int NS = kWavesPerWG * G;  // tile streams per workgroup
    constexpr int SROW = kTile * 2;      // float4 per tile of seg32
    constexpr int TROW = kTanRows * 2;   // float4 per tile of tangents
    __shared__ float4 lseg[NS][SROW + 1];
    __shared__ float4 ltan[NS][TROW + 1];
    __shared__ double red[kWavesPerWG][3][PB];
    const Work w = work[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int gl = lane % LG;
    const int stream = wave * G + lane / LG;
    double px[2], py[2], pz[2], ax[2], ay[2], az[2];
    float bx[2], by[2], bz[2], hbx[2], hby[2], hbz[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int i = gl + LG * j;
        const bool has = i < w.pt_count;
        const int64_t gp = (int64_t)w.pt_start + (has ? i : 0);
        px[j] = pts[gp];
        py[j] = pts[ptot + gp];
        pz[j] = pts[2 * ptot + gp];
        bx[j] = has ? (float)wts[gp] : 0.0f;  // surplus lanes carry no weight
        by[j] = has ? (float)wts[ptot + gp] : 0.0f;
        bz[j] = has ? (float)wts[2 * ptot + gp] : 0.0f;
        hbx[j] = 0.5f * bx[j];
        hby[j] = 0.5f * by[j];
        hbz[j] = 0.5f * bz[j];
        ax[j] = ay[j] = az[j] = 0.0;
    }
    const int tend = w.tile_start + w.tile_count;
    const float4* s4 = reinterpret_cast<const float4*>(seg32);
    const float4* t4 = reinterpret_cast<const float4*>(seg_tan);
    for (int t0 = w.tile_start; t0 < tend; t0 += NS) {
        const auto zero = [](int) { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); };
        stage_tiles<NS, SROW>(lseg, s4, t0, tend, zero);
        stage_tiles<NS, TROW>(ltan, t4, t0, tend, zero);
        __syncthreads();
        const int t = t0 + stream;
        if (t < tend) {  // uniform within a lane group
            const int cct = tile_cct[t];
            const float coef = cct >= 0 ? (float)cct_coef[cct] : 0.0f;
            float qx[2], qy[2], qz[2], sx[2], sy[2], sz[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                qx[j] = (float)(px[j] - tile_org[4 * t]);
                qy[j] = (float)(py[j] - tile_org[4 * t + 1]);
                qz[j] = (float)(pz[j] - tile_org[4 * t + 2]);
                sx[j] = sy[j] = sz[j] = 0.0f;
            }
            float cur = 0.0f;
            TanAcc ta{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            const float4* sr = lseg[stream];
            const float4* tr = ltan[stream];
#pragma unroll kBwdGrp2Unroll
            for (int k = 0; k < kTile; ++k) {
                const float4 kl = sr[2 * k], m = sr[2 * k + 1];
                float gkx, gky, gkz, ghx, ghy, ghz;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const BwdTerm b = bwd_term(qx[j], qy[j], qz[j], bx[j], by[j], bz[j], hbx[j], hby[j], hbz[j], kl, m);
                    sx[j] += b.hx;
                    sy[j] += b.hy;
                    sz[j] += b.hz;
                    cur += b.e;
                    if (j == 0) {
                        gkx = b.w * b.cx;
                        gky = b.w * b.cy;
                        gkz = b.w * b.cz;
                        ghx = b.hx;
                        ghy = b.hy;
                        ghz = b.hz;
                    } else {
                        gkx = __builtin_fmaf(b.w, b.cx, gkx);
                        gky = __builtin_fmaf(b.w, b.cy, gky);
                        gkz = __builtin_fmaf(b.w, b.cz, gkz);
                        ghx += b.hx;
                        ghy += b.hy;
                        ghz += b.hz;
                    }
                }
                const float4 e0 = tr[2 * k], e1 = tr[2 * k + 1];
                tan_step(ta, coef * e1.w, gkx, gky, gkz, ghx, ghy, ghz, e0, e1);
            }
            tan_end(ta, tr[2 * kTile], tr[2 * kTile + 1]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {  // s holds u / 2
                ax[j] += 2.0 * (double)sx[j];
                ay[j] += 2.0 * (double)sy[j];
                az[j] += 2.0 * (double)sz[j];
            }
            float g0 = ta.g0, g1 = ta.g1, g2 = ta.g2, g3 = ta.g3;
#pragma unroll
            for (int off = LG / 2; off > 0; off >>= 1) {  // sum over the group's lanes
                cur += __shfl_xor(cur, off);
                g0 += __shfl_xor(g0, off);
                g1 += __shfl_xor(g1, off);
                g2 += __shfl_xor(g2, off);
                g3 += __shfl_xor(g3, off);
            }
            if (gl == 0 && cct >= 0) {
                double* tg = tgrad + ((int64_t)w.block * ntile_total + t) * kTileGrad;
                tg[0] += (double)cur;
                tg[1] += (double)g0;
                tg[2] += (double)g1;
                tg[3] += (double)g2;
                tg[4] += (double)g3;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int off = LG; off < kWave; off <<= 1) {  // add the lane groups (fixed order)
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

// ------------------------------------------------------ stage adjoints
// Per-particle adjoint rows (adj[r][P]): 0-2 pbar, 3-5 vbar (end of step),
// 6-8 sum of P_k-bar, 9-11 sum of V_k-bar, 12-14 V_k-bar of the stage just done.
// bs_pts / bs_wts: points and weights of the next backward Biot-Savart pass.
#pragma clang fp contract(off)
__device__ inline void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// QS part of P-bar at point p with weight bb, and the QS gradients.
__device__ inline void qs_adjoint(int64_t d, const double* __restrict__ qs, const int64_t* __restrict__ design_qs_off,
                                  const double p[3], const double bb[3], double pbar[3], double* __restrict__ qgrad,
                                  int64_t ptot, int64_t g) {
    for (int64_t k = design_qs_off[d]; k < design_qs_off[d + 1]; ++k) {
        const double* q = qs + 16 * k;
        double x, y;
        if (!qs_local(q, p, x, y)) continue;
        const double gq = q[13], sq = q[14];
        const double bxi = bb[0] * q[3] + bb[1] * q[4] + bb[2] * q[5];
        const double byi = bb[0] * q[6] + bb[1] * q[7] + bb[2] * q[8];
        const double dx = bxi * (sq * y) + byi * (gq + sq * x);
        const double dy = bxi * (gq + sq * x) - byi * (sq * y);
        for (int c = 0; c < 3; ++c) pbar[c] = pbar[c] + (q[3 + c] * dx + q[6 + c] * dy);
        qgrad[(2 * k) * ptot + g] += bxi * y + byi * x;
        qgrad[(2 * k + 1) * ptot + g] += bxi * (x * y) + byi * (0.5 * (x * x - y * y));
    }
}

__device__ inline void load3(const double* __restrict__ a, int64_t ptot, int64_t g, double o[3]) {
    o[0] = a[g];
    o[1] = a[ptot + g];
    o[2] = a[2 * ptot + g];
}

__device__ inline void store3(double* __restrict__ a, int64_t ptot, int64_t g, const double v[3]) {
    a[g] = v[0];
    a[ptot + g] = v[1];
    a[2 * ptot + g] = v[2];
}

// A particle that has finished its steps takes no part in the backward pass: zero weight at a
// parking point away from every segment (pad rows of the segment set sit at +1e3, so the
// parking point must differ from them, else |r| = 0 there).
__device__ inline void set_bs(double* __restrict__ bs_pts, double* __restrict__ bs_wts, int64_t ptot, int64_t g,
                              const double p[3], const double w[3], bool active) {
    const double far[3] = {-1.0e3, -1.0e3, -1.0e3}, zero[3] = {0.0, 0.0, 0.0};
    store3(bs_pts, ptot, g, active ? p : far);
    store3(bs_wts, ptot, g, active ? w : zero);
}

// Stage 4 of the step: A4-bar = h/6 vbar, V4-bar = h/6 pbar + kq B4 x A4-bar, B4-bar = kq A4-bar x V4.
__global__ void k_adj_begin(const double* __restrict__ tape, double* __restrict__ adj, const double* __restrict__ kq,
                            const double* __restrict__ dt, const int64_t* __restrict__ nsteps, int64_t step,
                            int64_t ptot, double* __restrict__ bs_pts, double* __restrict__ bs_wts) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ptot) return;
    const bool active = step < nsteps[g];
    double pb[3], vb[3], P4[3], V4[3], B4[3];
    load3(adj, ptot, g, pb);
    load3(adj + 3 * ptot, ptot, g, vb);
    const double* t4 = tape + 3 * 9 * ptot;
    load3(t4, ptot, g, P4);
    load3(t4 + 3 * ptot, ptot, g, V4);
    load3(t4 + 6 * ptot, ptot, g, B4);
    const double h6 = dt[g] / 6, k = kq[g];
    double A[3], c[3], Vb[3], Bb[3];
    for (int i = 0; i < 3; ++i) A[i] = h6 * vb[i];
    cross3(B4, A, c);
    for (int i = 0; i < 3; ++i) Vb[i] = h6 * pb[i] + k * c[i];
    cross3(A, V4, c);
    for (int i = 0; i < 3; ++i) Bb[i] = k * c[i];
    const double zero[3] = {0.0, 0.0, 0.0};
    store3(adj + 6 * ptot, ptot, g, zero);
    store3(adj + 9 * ptot, ptot, g, Vb);
    store3(adj + 12 * ptot, ptot, g, Vb);
    set_bs(bs_pts, bs_wts, ptot, g, P4, Bb, active);
}

// After the backward pass at stage K+1 (1-based): P_{K+1}-bar, then stage K.
//   K=3: A3 = 2h/6 vbar + h V4b,   V3b = 2h/6 pbar + h P4b + kq B3 x A3
//   K=2: A2 = 2h/6 vbar + h/2 V3b, V2b = 2h/6 pbar + h/2 P3b + kq B2 x A2
//   K=1: A1 = h/6 vbar + h/2 V2b,  V1b = h/6 pbar + h/2 P2b + kq B1 x A1
//   K=0: only P1b; then pbar += sum P_k-bar, vbar += sum V_k-bar.
template <int K>
__global__ __launch_bounds__(kWG) void k_adj_stage(int ppl, const double* __restrict__ partial,
                                                   const int64_t* __restrict__ design_wbase,
                                                   const int32_t* __restrict__ design_nchunk,
                                                   const int64_t* __restrict__ design_pt_off, int64_t nd,
                                                   const double* __restrict__ qs,
                                                   const int64_t* __restrict__ design_qs_off,
                                                   const double* __restrict__ tape, double* __restrict__ adj,
                                                   const double* __restrict__ kq, const double* __restrict__ dt,
                                                   const int64_t* __restrict__ nsteps, int64_t step, int64_t ptot,
                                                   double* __restrict__ bs_pts, double* __restrict__ bs_wts,
                                                   double* __restrict__ qgrad) {
    const int64_t g = (int64_t)blockIdx.x * kWavesPerWG + (threadIdx.x >> 6);
    if (g >= ptot) return;
    const bool active = step < nsteps[g];
    double pb[3], vb[3], ps[3], vsum[3], vk[3], Pn[3], Bn[3];
    load3(adj, ptot, g, pb);
    load3(adj + 3 * ptot, ptot, g, vb);
    load3(adj + 6 * ptot, ptot, g, ps);
    load3(adj + 9 * ptot, ptot, g, vsum);
    load3(adj + 12 * ptot, ptot, g, vk);
    load3(bs_pts, ptot, g, Pn);  // P_{K+1} and its weight B_{K+1}-bar
    load3(bs_wts, ptot, g, Bn);
    const double h = dt[g], k = kq[g];
    const PointSlot sl = point_slot(g, ppl, partial, design_wbase, design_nchunk, design_pt_off, nd);
    if (!active) return;
    double Pb[3];
    chunk_sum(sl, Pb, adj);
    if ((threadIdx.x & (kWave - 1)) != 0) return;
    qs_adjoint(sl.design, qs, design_qs_off, Pn, Bn, Pb, qgrad, ptot, g);
    for (int i = 0; i < 3; ++i) ps[i] = ps[i] + Pb[i];
    if (K == 0) {
        for (int i = 0; i < 3; ++i) {
            pb[i] = pb[i] + ps[i];
            vb[i] = vb[i] + vsum[i];
        }
        store3(adj, ptot, g, pb);
        store3(adj + 3 * ptot, ptot, g, vb);
        return;
    }
    const double cv = K == 1 ? h / 6 : 2 * h / 6;   // weight of vbar/pbar in stage K
    const double cp = K == 3 ? h : h / 2;           // stage K+1 point = p + cp V_K
    const double* tk = tape + (K - 1) * 9 * ptot;
    double P[3], V[3], B[3];
    load3(tk, ptot, g, P);
    load3(tk + 3 * ptot, ptot, g, V);
    load3(tk + 6 * ptot, ptot, g, B);
    double A[3], c[3], Vb[3], Bb[3];
    for (int i = 0; i < 3; ++i) A[i] = cv * vb[i] + cp * vk[i];
    cross3(B, A, c);
    for (int i = 0; i < 3; ++i) Vb[i] = cv * pb[i] + cp * Pb[i] + k * c[i];
    cross3(A, V, c);
    for (int i = 0; i < 3; ++i) Bb[i] = k * c[i];
    for (int i = 0; i < 3; ++i) vsum[i] = vsum[i] + Vb[i];
    store3(adj + 6 * ptot, ptot, g, ps);
    store3(adj + 9 * ptot, ptot, g, vsum);
    store3(adj + 12 * ptot, ptot, g, Vb);
    set_bs(bs_pts, bs_wts, ptot, g, P, Bb, true);
}
#pragma clang fp contract(fast)

__global__ void k_copy_rows(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int PPL>
hipError_t launch_bwd(const float* seg32, const double* tile_org, const float* seg_tan, const int32_t* tile_cct,
                      const double* cct_coef, const Work* work, int64_t nwork, const double* pts, const double* wts,
                      int64_t ptot, int64_t ntile, double* partial, double* tgrad, hipStream_t st) {
    if (nwork == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bs_backward_f32<PPL>, dim3((unsigned)nwork), dim3(kWG), 0, st, seg32, tile_org, seg_tan,
                       tile_cct, cct_coef, work, pts, wts, ptot, ntile, partial, tgrad);
    return hipGetLastError();
}

hipError_t launch_bwd_any(int ppl, const float* seg32, const double* tile_org, const float* seg_tan,
                          const int32_t* tile_cct, const double* cct_coef, const Work* work, int64_t nwork,
                          const double* pts, const double* wts, int64_t ptot, int64_t ntile, double* partial,
                          double* tgrad, hipStream_t st) {
    if (ppl < 0) {
        if (nwork == 0) return hipSuccess;
        if (ppl == -2)  // 32-point blocks: 4 groups of 16 lanes, 2 points per lane
            hipLaunchKernelGGL(k_bs_backward_f32_grp2<4>, dim3((unsigned)nwork), dim3(kWG), 0, st, seg32, tile_org,
                               seg_tan, tile_cct, cct_coef, work, pts, wts, ptot, ntile, partial, tgrad);
        else
            hipLaunchKernelGGL(k_bs_backward_f32_grp<4>, dim3((unsigned)nwork), dim3(kWG), 0, st, seg32, tile_org,
                               seg_tan, tile_cct, cct_coef, work, pts, wts, ptot, ntile, partial, tgrad);
        return hipGetLastError();
    }
    switch (ppl) {
        case 1: return launch_bwd<1>(seg32, tile_org, seg_tan, tile_cct, cct_coef, work, nwork, pts, wts, ptot, ntile,
                                     partial, tgrad, st);
        case 2: return launch_bwd<2>(seg32, tile_org, seg_tan, tile_cct, cct_coef, work, nwork, pts, wts, ptot, ntile,
                                     partial, tgrad, st);
        default: return launch_bwd<4>(seg32, tile_org, seg_tan, tile_cct, cct_coef, work, nwork, pts, wts, ptot,
                                      ntile, partial, tgrad, st);
    }
}

}  // namespace
}  // namespace cctamd_detail

using namespace cctamd_detail;

extern "C" {

int cctamd_winding_tangents(const double* params, const int64_t* count_off, const int64_t* dst, int64_t ncct,
                            int64_t nseg_total, float* seg_tan, void* stream) {
    if (ncct < 0 || nseg_total < 0) return CCTAMD_EINVAL;
    if (ncct == 0 || nseg_total == 0) return CCTAMD_OK;
    if (!params || !count_off || !dst || !seg_tan) return CCTAMD_EINVAL;
    const int64_t npoint = nseg_total + ncct;  // count + 1 winding points per CCT
    const int64_t blocks = ceil_div(npoint, 256);
    if (blocks > 0x7fffffff) return CCTAMD_ERANGE;
    hipLaunchKernelGGL(k_winding_tangents, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, params,
                       count_off, dst, ncct, npoint, seg_tan);
    return hip_status(hipGetLastError());
}

int cctamd_track_rk4_taped(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                           const int32_t* work, int64_t nwork, const int64_t* design_wbase, const int32_t* design_nchunk,
                           const int64_t* design_pt_off, int64_t ndesign, const double* qs,
                           const int64_t* design_qs_off, const double* bconst, double* state, int64_t nparticles,
                           const double* kq, const double* dt, const int64_t* nsteps, int64_t max_steps,
                           double* scratch, double* partial, double* tape, void* stream) {
    if (!valid_common(precision, ppl, seg, tile_org, seg_near, nwork, ndesign)) return CCTAMD_EINVAL;
    if (nparticles < 0 || max_steps < 0) return CCTAMD_EINVAL;
    if (nparticles == 0 || max_steps == 0) return CCTAMD_OK;
    if (!state || !kq || !dt || !nsteps || !scratch || !tape || !design_wbase || !design_nchunk || !design_pt_off ||
        !design_qs_off || !bconst)
        return CCTAMD_EINVAL;
    if (nwork > 0 && (!work || !partial)) return CCTAMD_EINVAL;
    if (ceil_div(nparticles, kWavesPerWG) > 0x7fffffff) return CCTAMD_ERANGE;
    const DeviceSet ds{precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, design_wbase,
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, bconst};
    return hip_status(launch_rk4_steps(ds, state, nparticles, kq, dt, nsteps, 0, max_steps, scratch, partial, nullptr,
                                       1, tape, (hipStream_t)stream));
}

int cctamd_track_rk4_adjoint(int precision, int ppl, const void* seg, const double* tile_org,
                             const double* seg_near, const float* seg32, const float* seg_tan, const int32_t* tile_cct, const double* cct_coef, int64_t ntile,
                             const int32_t* work, int64_t nwork, const int64_t* design_wbase,
                             const int32_t* design_nchunk, const int64_t* design_pt_off, int64_t ndesign,
                             const double* qs, const int64_t* design_qs_off, const double* bconst,
                             const double* checkpoints, int64_t ckpt_every, int64_t nparticles, const double* kq,
                             const double* dt, const int64_t* nsteps, int64_t max_steps, double* adj, double* tgrad,
                             double* qgrad, double* state_ws, double* scratch, double* partial, double* tape,
                             double* adj_ws, double* bs_pts, double* bs_wts, void* stream) {
    if (!valid_common(precision, ppl, seg, tile_org, seg_near, nwork, ndesign)) return CCTAMD_EINVAL;
    if (nparticles < 0 || max_steps < 0 || ckpt_every < 0 || ntile < 0) return CCTAMD_EINVAL;
    if (nparticles == 0 || max_steps == 0) return CCTAMD_OK;
    const bool full_tape = ckpt_every == 0;  // tape of every step from cctamd_track_rk4_taped: no recompute
    if ((!full_tape && !checkpoints) || !kq || !dt || !nsteps || !adj || !tgrad || !qgrad || !state_ws || !scratch || !tape ||
        !adj_ws || !bs_pts || !bs_wts || !design_wbase || !design_nchunk || !design_pt_off || !design_qs_off ||
        !bconst)
        return CCTAMD_EINVAL;
    // the backward kernel reads the FP32 tiles and their origins whatever the forward precision
    if (nwork > 0 && (!work || !partial || !seg32 || !tile_org || !seg_tan || !tile_cct || !cct_coef))
        return CCTAMD_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const int64_t P = nparticles;
    const DeviceSet ds{precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, design_wbase,
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, bconst};
    const int64_t eblk = ceil_div(P, 256), rblk = ceil_div(P, kWavesPerWG);
    if (rblk > 0x7fffffff) return CCTAMD_ERANGE;
    hipError_t e;
    // adj rows 0-5 hold the seed; copy into the working rows
    hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)ceil_div(6 * P, 256)), dim3(256), 0, st, adj, adj_ws, 6 * P);
    const int64_t interval = full_tape ? max_steps : ckpt_every;
    const int64_t nck = ceil_div(max_steps, interval);
    for (int64_t c = nck - 1; c >= 0; --c) {
        const int64_t s0 = c * interval;
        const int64_t s1 = s0 + interval < max_steps ? s0 + interval : max_steps;
        if (!full_tape) {  // recompute this interval from its checkpoint, recording the tape
            hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)ceil_div(6 * P, 256)), dim3(256), 0, st,
                               checkpoints + c * 6 * P, state_ws, 6 * P);
            e = launch_rk4_steps(ds, state_ws, P, kq, dt, nsteps, s0, s1, scratch, partial, nullptr, 1, tape, st);
            if (e != hipSuccess) return (int)e;
        }
        for (int64_t step = s1 - 1; step >= s0; --step) {
            const double* tp = tape + (step - s0) * 36 * P;
            hipLaunchKernelGGL(k_adj_begin, dim3((unsigned)eblk), dim3(256), 0, st, tp, adj_ws, kq, dt, nsteps, step,
                               P, bs_pts, bs_wts);
#define CCTAMD_ADJ(K)                                                                                            \
    e = launch_bwd_any(ppl, seg32, tile_org, seg_tan, tile_cct, cct_coef, (const Work*)work, nwork, bs_pts,      \
                       bs_wts, P, ntile, partial, tgrad, st);                                                     \
    if (e != hipSuccess) return (int)e;                                                                          \
    hipLaunchKernelGGL(k_adj_stage<K>, dim3((unsigned)rblk), dim3(kWG), 0, st, ppl, partial, design_wbase,       \
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, tp, adj_ws, kq, dt, nsteps, step, \
                       P, bs_pts, bs_wts, qgrad);                                                                \
    e = hipGetLastError();                                                                                       \
    if (e != hipSuccess) return (int)e;
            CCTAMD_ADJ(3)
            CCTAMD_ADJ(2)
            CCTAMD_ADJ(1)
            CCTAMD_ADJ(0)
#undef CCTAMD_ADJ
        }
    }
    hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)ceil_div(6 * P, 256)), dim3(256), 0, st, adj_ws, adj, 6 * P);
    return hip_status(hipGetLastError());
}

}  // extern "C"
