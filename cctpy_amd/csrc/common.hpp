// Device helpers shared by the forward (cctamd.hip) and adjoint (adjoint.hip)
// kernels of libcctamd, and the host-side launch helpers they share.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cctamd.h"

namespace cctamd_detail {

constexpr int kTile = CCTAMD_TILE;
constexpr int kWave = 64;
constexpr int kWavesPerWG = 4;
constexpr int kWG = kWave * kWavesPerWG;
constexpr int64_t kXcds = 8;  // accelerator complex dies of an MI355X (each with its own L2)

inline int hip_status(hipError_t e) { return e == hipSuccess ? CCTAMD_OK : (int)e; }

// Points per block for a points-per-lane code: 1, 2, 4 -> 64, 128, 256 points;
// -2, -4 -> 32, 16 points (each wave split into 2 or 4 lane groups that walk
// different tiles with the same points).
__host__ __device__ inline int pb_of(int ppl) { return ppl > 0 ? kWave * ppl : kWave / (-ppl); }
inline bool valid_ppl(int ppl) { return ppl == 1 || ppl == 2 || ppl == 4 || ppl == -2 || ppl == -4; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------- row broadcasts
// Value of lane S of this 16-lane row, in every lane of the row (DPP
// row_newbcast, a move).
template <int S>
__device__ __forceinline__ float row_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + S, 0xf, 0xf,
                                                                  true));
}

// The helpers below put the broadcast on the instruction that consumes it: VOP2
// forms whose first source is read through the row_newbcast DPP control.  The
// compiler does not fold a DPP move into a multiply-accumulate, so a separate
// move per broadcast value would cost an instruction.  Their broadcast sources
// must not be written by a VALU instruction shortly before (the DPP hazard the
// compiler does not check in inline assembly; tests/test_dpp_hazards.py scans
// the built kernels).
// acc += (lane S of this 16-lane row's src) * mul.
template <int S>
__device__ __forceinline__ void fmac_bcast(float& acc, float src, float mul) {
    asm volatile("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "+v"(acc)
                 : "v"(src), "v"(mul), "i"(S));
}

// lane S of this row's src + x.
template <int S>
__device__ __forceinline__ float add_bcast(float src, float x) {
    float r;
    asm volatile("v_add_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "=v"(r)
                 : "v"(src), "v"(x), "i"(S));
    return r;
}

// Copies the rows of NS consecutive tiles (R float4 per tile, tiles t0 ..
// t0+NS-1 of src) into dst[NS][R + 1] in LDS; rows of tiles at or past tend
// get fill(e).  Each thread of the kWG-thread workgroup issues its global loads
// in batches of 4 before the batch's LDS writes, so the loads of a batch overlap
// instead of paying one memory round trip each.  The caller synchronises after.
template <int NS, int R, typename Fill>
__device__ __forceinline__ void stage_tiles(float4 (*dst)[R + 1], const float4* __restrict__ src, int t0, int tend,
                                            Fill fill) {
    constexpr int N = NS * R;
    constexpr int IT = (N + kWG - 1) / kWG;
    constexpr int BATCH = 4;
#pragma unroll 1
    for (int r0 = 0; r0 < IT; r0 += BATCH) {
        float4 v[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int i = (int)threadIdx.x + (r0 + b) * kWG;
            const int s = i / R, e = i % R;
            v[b] = fill(e);
            if (r0 + b < IT && i < N && t0 + s < tend) v[b] = src[(int64_t)(t0 + s) * R + e];
        }
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int i = (int)threadIdx.x + (r0 + b) * kWG;
            if (r0 + b < IT && i < N) dst[i / R][i % R] = v[b];
        }
    }
}

// ------------------------------------------------------------ geometry
// Mirrors CCT.__init__/phi_ksi_function/BipolarToroidalCoordinateSystem.convert
// (final_code/packages/cct.py:96-164, 166-188, 213-224) with the same
// operation order; contraction is off so each product is rounded like Python's.
#pragma clang fp contract(off)
__device__ inline double cct_phi(const double* __restrict__ prm, double ksi) {
    double phi = prm[17] * ksi + prm[18];
    const int ntilt = (int)prm[20];
    const int skip = (int)prm[29];
    for (int j = 0; j < ntilt && j < CCTAMD_MAX_TILTS; ++j) {
        if ((skip >> j) & 1) continue;
        phi = phi + prm[21 + j] * sin((double)(j + 1) * ksi);
    }
    return phi;
}

__device__ inline void cct_point(const double* __restrict__ prm, int64_t i, double out[3]) {
    const double ksi = prm[15] + prm[16] * (double)i;
    const double phi = cct_phi(prm, ksi);
    const double t = prm[12] / (prm[14] - cos(ksi));
    out[0] = t * prm[13] * cos(phi);
    out[1] = t * prm[13] * sin(phi);
    out[2] = t * sin(ksi);
}

// Hard-edge QS field (final_code/packages/magnets.py:637-668), same comparisons.
// Returns false outside the magnet; otherwise writes the local coordinates
// (x, y) and adds the field to b.
__device__ inline bool qs_local(const double* __restrict__ q, const double p[3], double& x, double& y) {
    const double d0 = p[0] + (-q[0]), d1 = p[1] + (-q[1]), d2 = p[2] + (-q[2]);
    x = q[3] * d0 + q[4] * d1 + q[5] * d2;
    y = q[6] * d0 + q[7] * d1 + q[8] * d2;
    const double z = q[9] * d0 + q[10] * d1 + q[11] * d2;
    const double L = q[12], a = q[15];
    if (z < 0 || z > L) return false;
    if (fabs(x) > a || fabs(y) > a || sqrt(x * x + y * y) > a) return false;
    return true;
}

__device__ inline void qs_add(const double* __restrict__ q, const double p[3], double b[3]) {
    double x, y;
    if (!qs_local(q, p, x, y)) return;
    const double g = q[13], s = q[14];
    const double bx = g * y + s * (x * y);
    const double by = g * x + 0.5 * s * (x * x - y * y);
    b[0] = b[0] + (q[3] * bx + q[6] * by);
    b[1] = b[1] + (q[4] * bx + q[7] * by);
    b[2] = b[2] + (q[5] * bx + q[8] * by);
}
#pragma clang fp contract(fast)

// ----------------------------------------------------- work decomposition
// One work item: a point block of one design x one chunk of its tiles.
// Partials of design d live at design_wbase[d]*3*PB and are laid out
// [block][point][component][chunk], so that the reduction over chunks reads
// contiguous memory.
struct Work {
    int32_t design, pt_start, pt_count, tile_start, tile_count, chunk, nchunk, block;
};

__device__ inline int64_t partial_base(const Work& w, int pb) {
    const int64_t wbase = (int64_t)blockIdx.x - (int64_t)w.block * w.nchunk - w.chunk;
    return wbase * 3 * pb + (int64_t)w.block * pb * 3 * w.nchunk + w.chunk;
}

template <int PPL>
__device__ inline void load_points(const double* __restrict__ pts, int64_t ptot, int pt_start, int pt_count,
                                   int lane, double px[PPL], double py[PPL], double pz[PPL]) {
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        int idx = lane + kWave * j;
        idx = idx < pt_count ? idx : 0;  // surplus lanes duplicate point 0, results discarded
        const int64_t g = (int64_t)pt_start + idx;
        px[j] = pts[g];
        py[j] = pts[ptot + g];
        pz[j] = pts[2 * ptot + g];
    }
}

// The NW waves of the workgroup hold partial sums for the same point block;
// they are added in wave order (deterministic) and written as one partial
// per (point, component): element (i, c) at out[(i*3 + c) * nchunk].
template <int PPL, int NW = kWavesPerWG>
__device__ inline void write_partial(double ax[PPL], double ay[PPL], double az[PPL], int wave, int lane,
                                     double* __restrict__ out, int nchunk) {
    constexpr int PB = kWave * PPL;
    __shared__ double red[NW][3][PB];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        red[wave][0][lane + kWave * j] = ax[j];
        red[wave][1][lane + kWave * j] = ay[j];
        red[wave][2][lane + kWave * j] = az[j];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * PB; k += kWave * NW) {
        const int c = k / PB, i = k % PB;
        double v = red[0][c][i];
#pragma unroll
        for (int q = 1; q < NW; ++q) v += red[q][c][i];
        out[(int64_t)(i * 3 + c) * nchunk] = v;
    }
}

__device__ inline int64_t find_design(const int64_t* __restrict__ off, int64_t nd, int64_t g) {
    int64_t lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Location of point g's partials: design, then the [component][chunk] rows.
struct PointSlot {
    int64_t design;
    const double* rows;  // rows[c * nc + chunk] for component c
    int nc;
};

__device__ inline PointSlot point_slot(int64_t g, int ppl, const double* __restrict__ partial,
                                       const int64_t* __restrict__ design_wbase,
                                       const int32_t* __restrict__ design_nchunk,
                                       const int64_t* __restrict__ design_pt_off, int64_t nd) {
    PointSlot s;
    s.design = find_design(design_pt_off, nd, g);
    const int pb = pb_of(ppl);
    const int64_t lp = g - design_pt_off[s.design];
    s.nc = design_nchunk[s.design];
    s.rows = partial + design_wbase[s.design] * 3 * pb + lp * 3 * (int64_t)s.nc;
    return s;
}

// The same slot precomputed per particle: scratch rows 12 (offset of the
// particle's partial rows, int64 bits) and 13 (nchunk | design << 32).
__device__ inline void store_slot(double* __restrict__ scratch, int64_t ptot, int64_t g, const PointSlot& s,
                                  const double* __restrict__ partial) {
    scratch[12 * ptot + g] = __longlong_as_double((long long)(s.rows - partial));
    scratch[13 * ptot + g] = __longlong_as_double(((long long)s.design << 32) | (long long)(unsigned)s.nc);
}

__device__ inline PointSlot load_slot(const double* __restrict__ scratch, int64_t ptot, int64_t g,
                                      const double* __restrict__ partial) {
    PointSlot s;
    const long long off = __double_as_longlong(scratch[12 * ptot + g]);
    const long long dn = __double_as_longlong(scratch[13 * ptot + g]);
    s.rows = partial + off;
    s.nc = (int)(unsigned)(dn & 0xffffffffll);
    s.design = dn >> 32;
    return s;
}

// Fixed-order sum over chunks (lanes stride the chunks, then a butterfly):
// deterministic, and every lane ends with the total.  A lane's loads for four
// of its chunks are issued together before any is added, the first four
// outside any loop so that they go out with the caller's other loads: one
// memory round trip for up to 256 chunks instead of one per 64.  The additions
// keep the chunk order.
template <int U>
__device__ __forceinline__ void chunk_batch(const double* __restrict__ rows, int nc, int c0, double& s0, double& s1,
                                            double& s2) {
    double v0[U], v1[U], v2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = c0 + u * kWave < nc ? c0 + u * kWave : (nc > 0 ? nc - 1 : 0);  // beyond nc: not added
        v0[u] = rows[c];
        v1[u] = rows[nc + c];
        v2[u] = rows[2 * nc + c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c0 + u * kWave < nc) {
            s0 += v0[u];
            s1 += v1[u];
            s2 += v2[u];
        }
    }
}

// readable: any valid device address, read (not used) in place of the rows of a
// design without segments (no chunks), so that the first loads need no branch.
// U chunk loads in flight per lane; the additions are in chunk order whatever U.
template <int U = 4>
__device__ inline void chunk_sum(const PointSlot& s, double b[3], const double* __restrict__ readable) {
    const int lane = threadIdx.x & (kWave - 1);
    const double* rows = s.nc > 0 ? s.rows : readable;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    chunk_batch<U>(rows, s.nc, lane, s0, s1, s2);
    for (int c0 = lane + U * kWave; c0 < s.nc; c0 += U * kWave) chunk_batch<U>(rows, s.nc, c0, s0, s1, s2);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    b[0] = s0;
    b[1] = s1;
    b[2] = s2;
}

// A design's analytic magnets: its QS range and constant field, loaded apart from
// their use so that a kernel can issue these loads next to its others.
struct AnalyticSet {
    int64_t k0, k1;
    double c0, c1, c2;
};

__device__ inline AnalyticSet analytic_set(int64_t d, const int64_t* __restrict__ design_qs_off,
                                           const double* __restrict__ bconst) {
    return AnalyticSet{design_qs_off[d], design_qs_off[d + 1], bconst[4 * d], bconst[4 * d + 1], bconst[4 * d + 2]};
}

__device__ inline void analytic_fields(const AnalyticSet& a, const double* __restrict__ qs, const double p[3],
                                       double b[3]) {
    for (int64_t k = a.k0; k < a.k1; ++k) qs_add(qs + 16 * k, p, b);
    b[0] = b[0] + a.c0;
    b[1] = b[1] + a.c1;
    b[2] = b[2] + a.c2;
}

__device__ inline void analytic_fields(int64_t d, const double* __restrict__ qs,
                                       const int64_t* __restrict__ design_qs_off,
                                       const double* __restrict__ bconst, const double p[3], double b[3]) {
    analytic_fields(analytic_set(d, design_qs_off, bconst), qs, p, b);
}

// ---------------------------------------------------- shared host helpers
struct DeviceSet {  // one segment set + analytic magnets of D designs
    int precision, ppl;
    const void* seg;
    const double* tile_org;
    const double* seg_near;  // FP64 rows for near tiles (precisions 32 and 33)
    const Work* work;
    int64_t nwork;
    const int64_t* design_wbase;
    const int32_t* design_nchunk;
    const int64_t* design_pt_off;
    int64_t ndesign;
    const double* qs;
    const int64_t* design_qs_off;
    const double* bconst;
};

bool valid_common(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                  int64_t nwork, int64_t ndesign);
hipError_t launch_bs_any(int precision, int ppl, const void* seg, const double* tile_org, const double* seg_near,
                         const Work* work, int64_t nwork, const double* pts, int64_t ptot, double* partial,
                         hipStream_t st);
// Checked field evaluation (cctamd_bs_field): Biot-Savart pass into `partial` (nwork * 3 * PB
// doubles), then the fixed-order reduction + QS + uniform fields into B. Returns a CCTAMD code.
int field_checked(const DeviceSet& ds, const double* points, int64_t npoints, double* partial, double* B,
                  hipStream_t st);
// RK4 steps [step0, step1) of every particle (forward), optional record of the
// state every `record_every` steps (row (step+1)/record_every) and optional
// tape of the stage points, velocities and fields relative to step0
// (tape[(step-step0)][stage][9][P]).
hipError_t launch_rk4_steps(const DeviceSet& ds, double* state, int64_t ptot, const double* kq, const double* dt,
                            const int64_t* nsteps, int64_t step0, int64_t step1, double* scratch, double* partial,
                            double* record, int64_t record_every, double* tape, hipStream_t st);

}  // namespace cctamd_detail
