/*
 * cctamd — C ABI of the MI355X (gfx950) Biot-Savart + RK4 tracking library.
 *
 * Every device pointer below is caller-owned device memory (the library never
 * allocates or frees it); every call is stream-ordered on `stream`
 * (a hipStream_t passed as void*, NULL = default stream) and returns
 * CCTAMD_OK, a negative CCTAMD_E* code for invalid arguments, or a positive
 * hipError_t.  No function keeps state between calls.
 *
 * Data layout in HBM (one "segment set" holds the CCT windings of D designs):
 *   seg64      double[S][8]  {klx, kly, klz, 0, mx, my, mz, 0}, global frame,
 *              kl = 1e-7*I*(P[i+1]-P[i]), m = (P[i+1]+P[i])/2; S is a multiple of
 *              CCTAMD_TILE; each design's range is padded with kl=0 segments.
 *   seg32      float[S][8]   {klx, kly, klz, 0, m-c, 0} with c = tile_org[tile]
 *   seg32k     float[S][12]  {kl, kl x (m-c), -2(m-c), |m-c|^2, 0, 0}
 *   tile_org   double[S/CCTAMD_TILE][4]  {c, e2}: FP64 tile origin (mean of the
 *              tile's current-carrying midpoints) and extent max |m-c|^2
 *   seg_near   the seg64 rows of the same set, passed beside the FP32 tiles: the FP32
 *              paths (precisions 32 and 33) sum a tile in FP64 from these rows when a
 *              point lies among its segments (|p-c|^2 < 1.5 e2), where FP32 terms would
 *              lose ~1e-6 of the nearest segment's share; ignored (may be NULL) at
 *              precision 64
 *   qs         double[Q][16] {O, XI, YI, ZI, length, gradient, second_gradient,
 *              aperture_radius} (the reference's QS record)
 *   bconst     double[D][4]  uniform field per design (sum of uniform magnets)
 *   points / particle state: SoA double[3][P] (x row, y row, z row)
 *   work       int32[W][8]   work items from cctamd_plan()
 */
#ifndef CCTAMD_H
#define CCTAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCTAMD_OK 0
#define CCTAMD_EINVAL (-22)
#define CCTAMD_ERANGE (-34)

#define CCTAMD_TILE 32          /* segments per tile */
#define CCTAMD_CCT_PARAMS 32    /* doubles per CCT record for cctamd_winding_generate */
#define CCTAMD_MAX_TILTS 8
#define CCTAMD_WORK_INTS 8

/* Library version (major*10000 + minor*100 + patch). */
int cctamd_version(void);

/* Message for a return code of this library. */
const char *cctamd_error_string(int code);

/*
 * Winding discretisation on the device.
 * Replaces CCT.__init__ path generation + global_current_elements_and_
 * elementary_current_positions (final_code/packages/cct.py:36-164, 166-188,
 * 213-224, 423-457), whose output the reference feeds to its GPU kernels
 * (final_code/packages/gpu_accelerator.py:1960-1966).
 *   params     double[ncct][CCTAMD_CCT_PARAMS] (device):
 *              [0..11] frame O, XI, YI, ZI; [12] a; [13] sinh(eta); [14] cosh(eta);
 *              [15] ksi0; [16] dksi = (ksi1-ksi0)*(1/(W*n)); [17] phi slope k;
 *              [18] phi intercept b; [19] 1e-7*current; [20] number of tilt terms;
 *              [21..28] tilt coefficients cot(t_j)/((j+1) sinh(eta)); [29] bit mask of
 *              tilt terms skipped (tilt == 90 deg); [30..31] unused.
 *   count_off  int64[ncct+1] (device): exclusive prefix sum of segment counts W*n
 *   dst        int64[ncct]   (device): first seg64 row of each CCT
 */
int cctamd_winding_generate(const double *params, const int64_t *count_off, const int64_t *dst,
                            int64_t ncct, int64_t nseg_total, double *seg64, void *stream);

/*
 * Path points of one CCT in its local frame (the reference's dispersed_path3,
 * final_code/packages/cct.py:134-148): out double[npoints][3], point i at
 * ksi = ksi0 + dksi*i, params as for cctamd_winding_generate (frame unused).
 */
int cctamd_winding_points(const double *params, int64_t npoints, double *out, void *stream);

/*
 * FP32 tile packing of a FP64 segment set (nseg a multiple of CCTAMD_TILE).
 * tile_org[t] = {c, e2}: c = mean of the tile's current-carrying midpoints (FP64),
 * e2 = max |m - c|^2 over them; seg32 rows are {kl, 0, m - c, 0} in FP32.
 * seg32k (nullable) receives the 12-float rows {kl, kl x (m-c), -2(m-c), |m-c|^2, 0, 0}
 * of precision 33 (see cctamd_bs_field).
 */
int cctamd_pack_fp32(const double *seg64, int64_t nseg, float *seg32, double *tile_org, float *seg32k,
                     void *stream);

/*
 * Host-only planner: splits (design, point block, tile chunk) work.
 *   design_ntile[d], design_npt[d], design_tile_off[d], design_pt_off[d] (host).
 *   ppl: points per lane (1, 2 or 4): a point block holds 64*ppl points; or -2 / -4:
 *        blocks of 32 / 16 points, each wave split into lane groups walking
 *        different tiles (precisions 33 and 64; not 32).
 * Every point block gets at least one item (one of no tiles for a design
 * without CCT tiles).
 * With work == NULL returns the number of work items only. Otherwise fills
 * work[nwork][8] = {design, pt_start, pt_count, tile_start, tile_count, chunk, nchunk, block},
 * design_wbase[d] (first item of design d) and design_nchunk[d].
 * Returns the number of items, or a negative code.
 */
int64_t cctamd_plan(int64_t ndesign, const int64_t *design_ntile, const int64_t *design_npt,
                    const int64_t *design_tile_off, const int64_t *design_pt_off, int ppl,
                    int64_t target_items, int64_t min_tiles, int32_t *work, int64_t *design_wbase,
                    int32_t *design_nchunk);

/*
 * Field of D designs at their points: B = sum over CCT segments + QS + uniform.
 * Replaces GPU_ACCELERATOR.current_element_B / magnet_at_qs_dates /
 * magnet_at_beamline_with_single_qs (final_code/packages/gpu_accelerator.py:1147-1196,
 * 1287-1347, 1349-1431) and CCT/Beamline.magnetic_field_at (cct.py:249-285,
 * beamline.py:46-53).
 *   precision: 64 (seg64, FP64 sum), 32 (seg32 + tile_org: FP32 terms of the
 *              re-centred tile, FP64 accumulation across tiles) or 33 (seg32k +
 *              tile_org: the same with per-tile sums K = sum w kl, C = sum w kl x m,
 *              B_tile = K x q - C and |q-m|^2 = (|q|^2 + |m|^2) - 2 q.m); at both FP32
 *              precisions tiles near a point are summed in FP64 from seg_near.
 *   seg, tile_org, seg_near: device pointers, required whenever nwork > 0 (a plan of
 *            cctamd_plan has an item for every point block, also when the designs have no
 *            segments: pass buffers of at least one row then; items without tiles read none).
 *   partial: device workspace of nwork*3*PB doubles (PB = points per block); design d's
 *            partials start at design_wbase[d]*3*PB, laid out [block][point][component][chunk].
 */
int cctamd_bs_field(int precision, int ppl, const void *seg, const double *tile_org, const double *seg_near,
                    const int32_t *work, int64_t nwork, const int64_t *design_wbase,
                    const int32_t *design_nchunk, const int64_t *design_pt_off, int64_t ndesign,
                    const double *qs, const int64_t *design_qs_off, const double *bconst,
                    const double *points, int64_t npoints, double *partial, double *B, void *stream);

/*
 * cctamd_bs_field without the partial workspace.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Arguments as for cctamd_bs_field
 * minus `partial`; the partial sums live in a static device array of
 * CCTAMD_NOALLOC_PARTIAL_DOUBLES doubles (16 MiB) that each GPU holds until the process
 * ends.  A plan needing more (nwork * 3 * PB doubles, PB = points per block) returns
 * CCTAMD_ERANGE before any launch.  Calls on one GPU share its array: calls that overlap
 * on one GPU (other streams or threads) race.  Allocates and frees nothing.
 */
#define CCTAMD_NOALLOC_PARTIAL_DOUBLES (2 * 1024 * 1024)
int cctamd_bs_field_noalloc(int precision, int ppl, const void *seg, const double *tile_org,
                            const double *seg_near, const int32_t *work, int64_t nwork,
                            const int64_t *design_wbase, const int32_t *design_nchunk,
                            const int64_t *design_pt_off, int64_t ndesign, const double *qs,
                            const int64_t *design_qs_off, const double *bconst, const double *points,
                            int64_t npoints, double *B, void *stream);

/*
 * The split-K Biot-Savart pass alone: for every work item, the sum over its
 * tiles for each point of its block, written in the partial layout of
 * cctamd_bs_field (no reduction).  The building block of
 * cctamd_bs_field / cctamd_track_rk4, exported for timing and for callers
 * that reduce partials themselves.
 */
int cctamd_bs_partial(int precision, int ppl, const void *seg, const double *tile_org, const double *seg_near,
                      const int32_t *work, int64_t nwork, const double *points, int64_t npoints,
                      double *partial, void *stream);

/*
 * RK4 tracking of D designs' particles through their own fields.
 * Replaces GPU_ACCELERATOR.track_multi_particle_beamline_for_magnet_with_multi_qs
 * (final_code/packages/gpu_accelerator.py:1912-2044) and the CPU ParticleRunner.run_only
 * (final_code/packages/particles.py:561-632 with base_utils.py:263-311).
 *   state    double[6][P] in/out: x, y, z, vx, vy, vz rows.
 *   scratch  double[15][P] device workspace.
 *   work     a plan of cctamd_plan (every point block has at least one item: with
 *            points-per-lane codes 1, 2, 4 each RK4 stage runs at the end of the
 *            Biot-Savart launch, by the last item of each point block to finish).
 *   kq[P] = e/m, dt[P], nsteps[P] (int64) per particle; steps beyond nsteps[i]
 *   leave particle i unchanged; max_steps = max(nsteps).
 *   record   nullable double[max_steps/record_every+1][6][P]: the state at step 0 and
 *            after every record_every-th step (row k = state after k*record_every
 *            steps; rows past a particle's own step count are left untouched).
 */
int cctamd_track_rk4(int precision, int ppl, const void *seg, const double *tile_org, const double *seg_near,
                     const int32_t *work, int64_t nwork, const int64_t *design_wbase,
                     const int32_t *design_nchunk, const int64_t *design_pt_off, int64_t ndesign,
                     const double *qs, const int64_t *design_qs_off, const double *bconst,
                     double *state, int64_t nparticles, const double *kq, const double *dt,
                     const int64_t *nsteps, int64_t max_steps, double *scratch, double *partial,
                     double *record, int64_t record_every, void *stream);

/*
 * Winding tangents for the adjoint, per tile of a segment set whose CCTs each start
 * on a tile boundary: float[ntile][33][8]; entry k < 32 is the winding point at the
 * start of the tile's segment k, entry 32 the end point of segment 31:
 * {dP/dphi(ksi).xyz, sin((j+1) ksi) j=0..3, real} in the global frame
 * (dP/dc_j = dP/dphi * sin((j+1) ksi), c_j the tilt coefficient of
 * final_code/packages/cct.py:179-187); real = 1 when segment k is a winding segment,
 * 0 for entry 32 and pad slots (the end point of a CCT's last segment also fills
 * the first pad slot of a partial tile).  The caller zero-fills seg_tan first.
 * Arguments as for cctamd_winding_generate (nseg_total = sum of the counts).
 */
int cctamd_winding_tangents(const double *params, const int64_t *count_off, const int64_t *dst,
                            int64_t ncct, int64_t nseg_total, float *seg_tan, void *stream);

/*
 * cctamd_track_rk4 (record_every = 1, no record) that also keeps the tape of every
 * step for cctamd_track_rk4_adjoint with ckpt_every = 0 (no counterpart in the
 * reference): tape double[max_steps][36][P] = per step and RK4 stage the stage
 * point, velocity and field rows.  Trades memory (288 B per particle-step) for the
 * recompute of the checkpointed mode.
 */
int cctamd_track_rk4_taped(int precision, int ppl, const void *seg, const double *tile_org,
                           const double *seg_near, const int32_t *work, int64_t nwork, const int64_t *design_wbase,
                           const int32_t *design_nchunk, const int64_t *design_pt_off, int64_t ndesign,
                           const double *qs, const int64_t *design_qs_off, const double *bconst,
                           double *state, int64_t nparticles, const double *kq, const double *dt,
                           const int64_t *nsteps, int64_t max_steps, double *scratch, double *partial,
                           double *tape, void *stream);

/*
 * Reverse-mode derivative of cctamd_track_rk4 (no counterpart in the reference,
 * which has no gradients; it differentiates the tracking that
 * final_code/packages/particles.py:561-632 performs).
 *   precision/seg: the forward path to recompute (32, 33 or 64);
 *   seg32, tile_org, seg_tan, tile_cct[ntile], cct_coef[ncct] (= 1e-7 I): FP32
 *     re-centred tiles of a segment set whose CCTs each start on a tile boundary;
 *     the backward kernel uses them at every precision, so tile_org (the FP32 tile
 *     origins of cctamd_pack_fp32) is required also when precision is 64;
 *   checkpoints: double[ceil(max_steps/ckpt_every)][6][P], the forward state every
 *     ckpt_every steps (cctamd_track_rk4 with record_every = ckpt_every); with
 *     ckpt_every = 0 there is no recompute: tape is the full tape written by
 *     cctamd_track_rk4_taped (double[max_steps][36][P]) and checkpoints may be NULL;
 *   adj: double[6][P], in: dJ/d(final x, v); out: dJ/d(initial x, v);
 *   tgrad: double[nblock][ntile][5] (+=): per tile sum B-bar.dB (current; divide by
 *     I) and dJ/dc_0..3 (tilt coefficients);
 *   qgrad: double[2*Q][P] (+=): dJ/d(gradient), dJ/d(second gradient) of each QS;
 *   workspaces: state_ws[6][P], scratch[15][P], partial (as cctamd_bs_field),
 *     tape[ckpt_every][36][P] (or the full tape), adj_ws[15][P], bs_pts[3][P], bs_wts[3][P].
 */
int cctamd_track_rk4_adjoint(int precision, int ppl, const void *seg, const double *tile_org,
                             const double *seg_near, const float *seg32, const float *seg_tan, const int32_t *tile_cct,
                             const double *cct_coef, int64_t ntile, const int32_t *work, int64_t nwork,
                             const int64_t *design_wbase, const int32_t *design_nchunk,
                             const int64_t *design_pt_off, int64_t ndesign, const double *qs,
                             const int64_t *design_qs_off, const double *bconst,
                             const double *checkpoints, int64_t ckpt_every, int64_t nparticles,
                             const double *kq, const double *dt, const int64_t *nsteps, int64_t max_steps,
                             double *adj, double *tgrad, double *qgrad, double *state_ws, double *scratch,
                             double *partial, double *tape, double *adj_ws, double *bs_pts, double *bs_wts,
                             void *stream);

/*
 * Phase-space projection of tracked particles and beam statistics per group.
 * Replaces PhaseSpaceParticle.create_from_running_particle(s)
 * (final_code/packages/particles.py:1077-1130) applied to every design of a
 * population, and the beam-size statistics of the design loop
 * (cctpy3/run.py:72-103: (max - min) / 2 of x and y per group of particles).
 *   ideal: double[ndesign][16] = {p.xyz, v.xyz, P (speed * relativistic mass),
 *          XI.xyz, YI.xyz, ZI.xyz} of the ideal particle and its natural frame;
 *   state: double[6][P] tracked states; speed[P], momentum[P]: each particle's
 *          scalar speed and speed * relativistic mass (RunningParticle fields);
 *   group_off[G+1] (particle ranges, clamped to [0, P]) and group_design[G]
 *          (index into ideal, clamped to [0, ndesign)), device arrays;
 *   pps: double[6][P] out, rows x, xp, y, yp, z, delta;
 *   stats: double[G][12] out = {n, min x, max x, min y, max y, mean x, mean y,
 *          mean delta, rms x, rms y, sum (delta - mean)^2, sum (x - mean x)(delta - mean delta)}.
 * Reductions run in a fixed order (deterministic).
 */
int cctamd_phase_space(const double *ideal, int64_t ndesign, const double *state, const double *speed,
                       const double *momentum, int64_t nparticles, const int64_t *group_off,
                       const int32_t *group_design, int64_t ngroups, double *pps, double *stats,
                       void *stream);

/*
 * Adjoint seed of the beam-spot objective J_g = w0 rms x + w1 rms y + w2 |D| of
 * each group (D = dispersion, stats[11] / stats[10]; no counterpart in the
 * reference, which has no gradients): writes dJ/d(final x, v) into adj[6][P] for
 * the particles of every group (the velocity rows are zero).  weights: host double[3].
 */
int cctamd_beam_objective_seed(const double *ideal, int64_t ndesign, const double *pps, const double *stats,
                               int64_t nparticles, const int64_t *group_off, const int32_t *group_design,
                               int64_t ngroups, const double *weights, double *adj, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CCTAMD_H */
