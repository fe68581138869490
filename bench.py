#!/usr/bin/env python3
"""Throughput of RK4 proton tracking through CCT gantry sections on one MI355X.

A *step* is one full tracking pass of the whole synthetic beam through the
workload's beamline (all RK4 steps; default workload ``agcct_section``:
368,640 winding segments + 2 quads, 512 protons, 2,510 RK4 steps of 1 mm).
The timed passes use the FP32 Biot-Savart tracking path (per-tile
re-centred FP32 terms summed as K x q - C per tile, FP64 accumulation across
tiles, FP64 particle state); one pass of the plain FP64 path is timed beside
it and is the reference for the final-position error.
Inputs are resident in HBM before timing starts; every pass restarts from the
same initial state with a device-to-device copy.

Besides the main line, ``extras`` (skip with ``--no-extras``) reports
  * the batched design population (64 designs of uneven segment counts, 32
    protons each): FP32 forward, plain FP64 forward, and FP32 forward+adjoint
    (beam-spot objective per design, checkpoints every 50 steps);
  * the differentiation workload (AG-CCT section + 1.1 m drift, 256 protons):
    FP32 forward+adjoint.

Prints one JSON line on stdout (diagnostics go to stderr).

Profiling helpers (not bench lines): ``--roofline-only`` launches only the
Biot-Savart kernel on the workload's initial points (for counter passes),
``--copy-mib M`` runs a plain elementwise copy of M MiB (the calibration of
the HBM byte counters).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 27           # r (3) + |r|^2 (5) + sqrt (1) + r^3 (2) + kl x r (9) + 1/x (1) + scale (3) + sum (3)
PEAK_TFLOPS = {32: 157.3, 64: 78.6}
METRIC = "particle-steps/s and B-S pair-evals/s vs plain FP64 GPU, at ≤1 µm position error"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bs_kernel_traffic.json")
FP32 = 33                    # FP32 tracking kernel (K x q - C tile sums)
CKPT_EVERY = 50


def log(*a):
    print(*a, file=sys.stderr, flush=True)


BS_WAVES = 8                 # waves per workgroup of k_bs_partial_f32k (CCTAMD_BS_WAVES in csrc/cctamd.hip)


def kernel_name(ppl: int) -> str:
    """Name of the FP32 tracking kernel a points-per-lane code launches (32-point blocks: rows of
    16 lanes fed by DPP row broadcasts; 16-point blocks: 8 lane groups fed from LDS)."""
    return f"k_bs_partial_f32k<{ppl}, {BS_WAVES}>" if ppl > 0 else (
        "k_bs_partial_f32k_row" if ppl == -2 else "k_bs_partial_f32k_grp<8, 2>")


def measured_traffic(workload: str, kernel: str, nwork: int):
    """HBM bytes per launch of this kernel shape from the committed counter summary, or None.  Entries
    carry the kernel-set version they were measured with (autotune.KERNEL_SET, raised with every
    kernel change): a summary of older kernels gives None."""
    from cctpy_amd import autotune
    try:
        with open(TRAFFIC_FILE) as f:
            rows = json.load(f)["launches"]
    except (OSError, ValueError, KeyError):
        return None
    for r in rows:
        if (r.get("kernel_set") == autotune.KERNEL_SET and r["workload"] == workload and r["kernel"] == kernel
                and r["work_items"] == nwork):
            return r["hbm_bytes_per_launch"]
    return None


def copy_calibration(mib: int, reps: int = 5) -> int:
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.neg float32", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "dispatches": reps}), flush=True)
    return 0


def cpu_baseline(workload: str, sample: str, workers: int = 16):
    """The reference CPU path as the oracle restates it, on a bounded sample of the workload, in a
    child process that never touches the GPU (oracle/cpu_baseline.py): the numpy restatement
    (per-particle RK4, numpy Biot-Savart per CCT, the reference's operation order) in one process
    and in a pool of `workers` processes (one task per proton, like
    ParticleRunner.run_only(concurrency_level=os.cpu_count())), and the C restatement (OpenMP
    over protons) on the pool's sample.  sample = "N1xS1,NPxSP": protons x RK4 steps of the
    one-process and of the pool runs.  Called before this process touches the GPU."""
    import subprocess
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    workers = max(1, min(workers, cores))
    one, pool = sample.split(",")
    n1, s1 = (int(v) for v in one.split("x"))
    np_, sp = (int(v) for v in pool.split("x"))
    res = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", workload, str(n1), str(s1), str(np_), str(sp),
                          str(workers)], cwd=ROOT, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + res.stderr[-2000:])
    r = json.loads(res.stdout.strip().splitlines()[-1])
    model = open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(" :\t") \
        if os.path.exists("/proc/cpuinfo") else "unknown"
    rate = r["numpy_pool"]["particle_steps_per_s"]
    pairs = 4 * r["segments"]
    return {"value": rate, "unit": "particle-steps/s", "cores": workers, "kind": "port",
            "os_cpu_count": os.cpu_count(), "cores_visible": cores, "cpu_model": model,
            "sample": f"{np_} protons x {sp} RK4 steps of {workload} over all {r['segments']} segments + QS: numpy "
                      f"restatement of the reference CPU path (per-particle RK4, numpy Biot-Savart per CCT), "
                      f"{workers} processes, one task per proton ({r['numpy_pool']['seconds']:.1f} s); host cores "
                      f"visible {cores}, CPU {model}",
            "pair_evals_per_s": rate * pairs,
            "numpy_1_process": {"particle_steps_per_s": r["numpy_1proc"]["particle_steps_per_s"],
                                "pair_evals_per_s": r["numpy_1proc"]["particle_steps_per_s"] * pairs,
                                "sample": f"{n1} protons x {s1} RK4 steps, one process"},
            "c_openmp": {"particle_steps_per_s": r["c_openmp"]["particle_steps_per_s"],
                         "pair_evals_per_s": r["c_openmp"]["particle_steps_per_s"] * pairs,
                         "threads": workers, "kind": "port",
                         "sample": f"C FP64 restatement, OpenMP {workers} threads, same {np_} x {sp} sample"}}


class Setup:
    """A workload resident on the device: designs, particles, step counts."""

    def __init__(self, wl, align_cct: bool = False, max_steps: int = 0):
        import numpy as np
        import torch
        from cctpy_amd import engine
        from cctpy_amd import workloads as W
        from cctpy_amd.magnets import device_parts
        self.wl = wl
        self.bls = [W.build_beamline(ops) for ops in wl.designs]
        ps, npt = [], []
        for bl, beam in zip(self.bls, wl.beams):
            q = W.beam_particles(beam, bl)
            ps += q
            npt.append(len(q))
        self.ps, self.npt, self.P = ps, npt, len(ps)
        state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
        kq = np.array([p.e / p.relativistic_mass for p in ps])
        lengths = [wl.beam_length(i) for i, n in enumerate(npt) for _ in range(n)]
        steps_dt = [engine.rk4_steps(L, wl.footstep, p.speed) for L, p in zip(lengths, ps)]
        self.ns = np.array([s for s, _ in steps_dt], np.int64)
        if max_steps > 0:  # a prefix of every track (same per-step work)
            self.ns = np.minimum(self.ns, max_steps)
        dt = np.array([d for _, d in steps_dt])
        self.max_steps = int(self.ns.max())
        t0 = time.perf_counter()
        self.model = engine.DeviceModel([device_parts(bl).spec() for bl in self.bls], fp32=True,
                                        align_cct=align_cct)
        torch.cuda.synchronize()
        self.gen_s = time.perf_counter() - t0
        dev = engine.device()
        self.s0 = torch.from_numpy(np.ascontiguousarray(state.T)).to(dev)
        self.kq = torch.from_numpy(kq).to(dev)
        self.dt = torch.from_numpy(dt).to(dev)
        self.ns_t = torch.from_numpy(self.ns).to(dev)
        self.st = torch.empty_like(self.s0)
        self.particle_steps = float(self.ns.sum())
        seg_of = np.repeat(self.model.design_nseg, npt)
        self.pairs_per_pass = float(np.sum(self.ns * 4 * seg_of))
        log(f"[bench] {wl.name}: {len(self.bls)} design(s), {self.model.nseg} segments, {self.P} protons, "
            f"{self.max_steps} RK4 steps; windings generated on the GPU in {self.gen_s:.2f} s")

    def track_samples(self, n: int = 8):
        """Positions (3, P) sampled along a quick FP32 track with a default shape."""
        import torch
        from cctpy_amd import engine
        plan = self.model.plan(self.npt, 2)
        scratch, partial = self.model.workspace(plan)
        every = max(1, self.max_steps // n)
        rec = torch.empty((self.max_steps // every + 1, 6, self.P), dtype=torch.float64, device=engine.device())
        self.st.copy_(self.s0)
        self.model.track_tensors(plan, FP32, self.st, self.kq, self.dt, self.ns_t, self.max_steps, scratch, partial,
                                 record=rec, record_every=every)
        return [rec[i, :3].contiguous() for i in range(rec.shape[0])]

    def shape(self, prec: int, ppl: int, items_per_cu: int) -> dict:
        from cctpy_amd import autotune, engine
        if ppl == 0:
            if not hasattr(self, "_samples"):
                self._samples = self.track_samples()
            return autotune.tune(self.model, self.npt, prec, self._samples)
        if prec == 32 and ppl < 0:
            ppl = 2
        return {"ppl": ppl, "target_items": items_per_cu * engine.cu_count() or None}

    def timed(self, one_pass, passes: int, warmup: int) -> float:
        import torch
        for _ in range(warmup):
            self.st.copy_(self.s0)
            one_pass()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(passes):
            self.st.copy_(self.s0)
            one_pass()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    def forward(self, prec: int, shape: dict, passes: int, warmup: int, graph: bool):
        """(ms per pass, final state (P, 6), plan, partial) of the tracking path at one precision."""
        from cctpy_amd import engine
        m = self.model
        plan = m.plan(self.npt, shape["ppl"], shape["target_items"])
        scratch, partial = m.workspace(plan)
        if graph:
            g = engine.TrackGraph(m, plan, prec, self.st, self.kq, self.dt, self.ns_t, self.max_steps, scratch,
                                  partial)
            one_pass = g.replay
        else:
            def one_pass():
                m.track_tensors(plan, prec, self.st, self.kq, self.dt, self.ns_t, self.max_steps, scratch, partial)
        ms = self.timed(one_pass, passes, warmup) / passes
        return ms, self.st.cpu().numpy().T.copy(), plan, partial

    def forward_adjoint(self, shape: dict, passes: int, warmup: int, ckpt_every: int = CKPT_EVERY,
                        tune: bool = True):
        """ms per FP32 forward (with checkpoints every ckpt_every steps, or the full tape with 0)
        + objective + adjoint pass, and the split.  With ``tune`` the block shape is the one
        with the fastest short forward + adjoint pass; otherwise ``shape``'s."""
        import torch
        from cctpy_amd import engine
        from cctpy_amd.adjoint import AdjointTracker
        from cctpy_amd.beam import BeamGroups
        from cctpy_amd.particles import ParticleFactory
        ipc = shape["target_items"] // engine.cu_count() if shape.get("target_items") else None
        tr = AdjointTracker(self.bls, precision=FP32, ckpt_every=ckpt_every, ppl=shape["ppl"], items_per_cu=ipc)
        if tune:  # forward + backward decide the block shape, not the forward kernel alone
            tr.tune(self.npt, self.s0, self.kq, self.dt, self.ns_t)
        tr.prepare(self.npt, self.kq, self.dt, self.ns_t, self.max_steps)
        ideals, frames = [], []
        for i, (bl, beam) in enumerate(zip(self.bls, self.wl.beams)):
            ip = ParticleFactory.create_proton_along(bl, beam.s0 + self.wl.beam_length(i), beam.kinetic_mev)
            ideals.append(ip)
            frames.append(ip.get_natural_coordinate_system())
        bg = BeamGroups(ideals, frames, self.npt, [p.speed for p in self.ps],
                        [p.compute_scalar_momentum() for p in self.ps])
        out = {}

        def one_pass():
            tr.forward_tensors(self.st)
            pps, stats = bg.project(self.st)
            adj = bg.seed(pps, stats)
            tr.backward_tensors(adj)
            out["stats"] = stats

        ms = self.timed(one_pass, passes, warmup) / passes
        fwd = self.timed(lambda: tr.forward_tensors(self.st), 1, 0)
        J = BeamGroups.objective(out["stats"])
        return {"ms_per_pass": ms, "forward_ms": fwd, "adjoint_ms": ms - fwd,
                "particle_steps_per_s": self.particle_steps / (ms / 1e3), "ppl": tr.plan.ppl,
                "items_per_cu": tr.items_per_cu,
                "checkpoint_every": ckpt_every, "objective_mean": float(J.mean()),
                "grad_finite": bool(torch.isfinite(tr.tgrad).all().item())}


def tracking_leg(name: str, max_steps: int = 0, passes: int = 3, ppl: int = 0, items_per_cu: int = 0,
                 adjoint_steps: int = 0) -> dict:
    """FP32 and plain FP64 tracking passes of one workload (HIP-graph replay, inputs resident),
    their final-position difference, and optionally FP32 forward + adjoint over a step prefix.
    ppl = 0 lets the autotuner pick the block shape (it tracks a sample pass first)."""
    import numpy as np
    from cctpy_amd import workloads as W
    S = Setup(W.by_name(name), align_cct=adjoint_steps > 0, max_steps=max_steps)
    sh32, sh64 = S.shape(FP32, ppl, items_per_cu), S.shape(64, ppl, items_per_cu)
    ms32, out32, plan32, _ = S.forward(FP32, sh32, passes, 1, True)
    ms64, out64, _, _ = S.forward(64, sh64, max(1, passes // 2), 1, True)
    d = np.linalg.norm(out32[:, :3] - out64[:, :3], axis=1)
    res = {"segments": int(S.model.nseg), "protons": S.P, "rk4_steps": S.max_steps, "steps_capped": max_steps > 0,
           "footstep_mm": S.wl.footstep * 1e3, "hip_graph": True,
           "fp32": {"ms_per_pass": ms32, "particle_steps_per_s": S.particle_steps / (ms32 / 1e3),
                    "pair_evals_per_s": S.pairs_per_pass / (ms32 / 1e3), "shape": sh32},
           "fp64": {"ms_per_pass": ms64, "particle_steps_per_s": S.particle_steps / (ms64 / 1e3),
                    "pair_evals_per_s": S.pairs_per_pass / (ms64 / 1e3), "shape": sh64},
           "fp32_vs_fp64": ms64 / ms32,
           "launches_per_rk4_step": 8,
           "position_error_um_vs_fp64": {"max": float(d.max() * 1e6), "median": float(np.median(d) * 1e6),
                                         "p99": float(np.percentile(d, 99) * 1e6),
                                         "within_1um": float(np.mean(d <= 1e-6))}}
    if adjoint_steps > 0:
        A_ = Setup(W.by_name(name), align_cct=True, max_steps=adjoint_steps)
        res["forward_adjoint"] = dict(rk4_steps=A_.max_steps, **A_.forward_adjoint(
            A_.shape(FP32, 2, 4), 1, 0, ckpt_every=0, tune=False))
        del A_
    log(f"[bench] {name}: fp32 {ms32:.1f} ms, fp64 {ms64:.1f} ms per pass of {S.max_steps} steps, error max "
        f"{d.max() * 1e6:.3g} um")
    return res


def field_along_arc(n_points: int = 1000, reps: int = 20) -> dict:
    """Field of the single-CCT workload (23,040 segments) at points spaced evenly along the ideal
    arc in the median plane: FP64 and FP32 paths timed (HIP events, inputs on the device), the
    FP32 relative error against FP64 over points whose field is at least 1 % of the largest."""
    import numpy as np
    import torch
    from cctpy_amd import engine
    from cctpy_amd import workloads as W
    from cctpy_amd.magnets import device_parts
    wl = W.single_cct(n_particles=1)
    bl = W.build_beamline(wl.designs[0])
    model = engine.DeviceModel([device_parts(bl).spec()], fp32=True)
    s0 = wl.beams[0].s0
    pts = np.array([[q.x, q.y, 0.0] for q in (bl.point_at(float(v))
                                                 for v in np.linspace(s0, s0 + wl.length, n_points))])
    dev = engine.device()
    soa = torch.from_numpy(np.ascontiguousarray(pts.T)).to(dev)
    out, res = {}, {}
    for prec, ppl in ((64, 1), (FP32, 1)):
        plan = model.plan([n_points], ppl)
        _, partial = model.workspace(plan)
        b = torch.empty_like(soa)
        model.field_tensors(plan, prec, soa, partial, b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            model.field_tensors(plan, prec, soa, partial, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[prec] = b.clone()
        out["fp64" if prec == 64 else "fp32"] = {"ms": ms, "pair_evals_per_s": n_points * model.nseg / (ms / 1e3)}
    dn, bn = (res[FP32] - res[64]).norm(dim=0), res[64].norm(dim=0)
    keep = bn >= 0.01 * bn.max()
    out.update({"points": n_points, "segments": int(model.nseg), "launches_per_evaluation": 2,
                "fp32_max_rel_err": float((dn[keep] / bn[keep]).max()), "peak_T": float(bn.max())})
    return out


def gradient_check(n_particles: int = 16) -> dict:
    """dJ/dtheta by the adjoint (FP32 Biot-Savart, full tape) for a linear functional of the final
    states, over the 10-parameter design vector of the adjoint_section workload (full geometry and
    track, fewer protons), against central differences (h = 1e-6 relative) of the FP64 forward."""
    import numpy as np
    from cctpy_amd import adjoint
    from cctpy_amd import workloads as W
    wl = W.adjoint_section(n_particles=n_particles)
    beam, length = wl.beams[0], wl.beam_length(0)

    def build(th):
        return W.build_beamline(W.section_ops(th))

    def particles(bl):
        return W.beam_particles(beam, bl)

    theta = W.section_theta0()
    t0 = time.perf_counter()
    prob = adjoint.DesignProblem(build, particles, length, wl.footstep, precision=FP32, ckpt_every=0)
    final = prob.forward(theta)
    w = np.random.default_rng(11).normal(size=final.shape) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    grad = prob.backward(w)
    fd_prob = adjoint.DesignProblem(build, particles, length, wl.footstep, precision=64, ckpt_every=0)
    fd = np.zeros_like(theta)
    for k in range(theta.size):
        h = 1e-6 * max(abs(theta[k]), 1.0)
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (float(np.sum(w * fd_prob.forward(tp))) - float(np.sum(w * fd_prob.forward(tm)))) / (2 * h)
    # per parameter: |g_k - fd_k| / max(|fd_k|, 1e-3 max |fd|) (the parameters mix T/m, degrees
    # and amperes, so one figure normalised by the largest component would hide a small one)
    floor = 1e-3 * np.abs(fd).max()
    rel = np.abs(grad - fd) / np.maximum(np.abs(fd), floor)
    return {"parameters": int(theta.size), "names": list(W.THETA_NAMES), "protons": int(final.shape[0]),
            "segments": int(prob.tracker.model.nseg),
            "functional": "w . final state (seeded Gaussian w)", "fd_step": "1e-6 relative, FP64 forward",
            "error_definition": "|g_k - fd_k| / max(|fd_k|, 1e-3 max_k |fd_k|)",
            "max_rel_err_vs_fd": float(rel.max()), "rel_err_per_parameter": [float(v) for v in rel],
            "fd_per_parameter": [float(v) for v in fd], "adjoint_per_parameter": [float(v) for v in grad],
            "seconds": time.perf_counter() - t0}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed tracking passes")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="agcct_section")
    ap.add_argument("--ppl", type=int, default=0, help="points per lane, -2/-4 split waves (0: autotune)")
    ap.add_argument("--items-per-cu", type=int, default=0, help="work items per CU with --ppl (0: default)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison pass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip the population and adjoint extras")
    ap.add_argument("--cpu-sample", default="4x12,32x12",
                    help="CPU baseline samples: protons x RK4 steps of the one-process and of the pool runs")
    ap.add_argument("--roofline-launches", type=int, default=20)
    ap.add_argument("--roofline-precision", type=int, choices=(FP32, 64), default=FP32,
                    help="with --roofline-only: launch the FP32 (33) or the plain FP64 (64) kernel")
    ap.add_argument("--roofline-only", action="store_true", help="only launch the Biot-Savart kernel (profiling)")
    ap.add_argument("--copy-mib", type=int, default=0, help="plain copy of this many MiB (counter calibration)")
    ap.add_argument("--graph", action="store_true", help="replay each tracking pass as one captured HIP graph")
    ap.add_argument("--max-steps", type=int, default=0, help="cap every track at this many RK4 steps (0: full)")
    ap.add_argument("--ckpt-every", type=int, default=CKPT_EVERY, help="adjoint mode: checkpoint interval (0: full tape)")
    ap.add_argument("--mode", choices=("forward", "adjoint"), default="forward",
                    help="adjoint: time only FP32 forward + objective + adjoint passes of the workload (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_mib:
        return copy_calibration(args.copy_mib)

    # the CPU baselines first, in child processes, before this process touches the GPU
    cpu = cpu_cfg1 = None
    if not (args.no_cpu or args.roofline_only or args.mode == "adjoint"):
        cpu = cpu_baseline(args.workload, args.cpu_sample)
        log(f"[bench] cpu baseline: numpy {cpu['value']:.4g} particle-steps/s ({cpu['cores']} processes), "
            f"{cpu['numpy_1_process']['particle_steps_per_s']:.3g} in one process; C/OpenMP "
            f"{cpu['c_openmp']['particle_steps_per_s']:.4g}")
        if not args.no_extras:  # the single CCT's whole run (64 protons x 498 steps), not a slice
            cpu_cfg1 = cpu_baseline("single_cct", "1x498,64x498")
            log(f"[bench] cpu single_cct in full: numpy {cpu_cfg1['value']:.4g} particle-steps/s "
                f"({cpu_cfg1['cores']} processes); C/OpenMP {cpu_cfg1['c_openmp']['particle_steps_per_s']:.4g}")

    import numpy as np
    import torch

    from cctpy_amd import engine
    from cctpy_amd import workloads as W

    if args.mode == "adjoint":
        A_ = Setup(W.by_name(args.workload), align_cct=True, max_steps=args.max_steps)
        fa = A_.forward_adjoint(A_.shape(FP32, args.ppl, args.items_per_cu), args.steps, args.warmup,
                                ckpt_every=args.ckpt_every, tune=args.ppl == 0)
        print(json.dumps(dict(workload=A_.wl.name, protons=A_.P, segments=int(A_.model.nseg),
                              rk4_steps=A_.max_steps, **fa)), flush=True)
        return 0

    S = Setup(W.by_name(args.workload), max_steps=args.max_steps)
    wl, model = S.wl, S.model
    tuned = {FP32: S.shape(FP32, args.ppl, args.items_per_cu)}
    log(f"[bench] fp32 shape: {tuned[FP32]}")

    def bs_launch_timing(plan, pts, partial, launches: int, prec: int = FP32) -> float:
        """Average duration (ms) of the Biot-Savart kernel, HIP events on its stream."""
        model.bs_partial_tensors(plan, prec, pts, partial)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(launches):
            model.bs_partial_tensors(plan, prec, pts, partial)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / launches

    if args.roofline_only:
        plan = model.plan(S.npt, tuned[FP32]["ppl"], tuned[FP32]["target_items"])
        _, partial = model.workspace(plan)
        avg = bs_launch_timing(plan, S.s0[:3].contiguous(), partial, args.roofline_launches, args.roofline_precision)
        kname = kernel_name(plan.ppl) if args.roofline_precision == FP32 else (
            f"k_bs_partial_f64<{plan.ppl}, {BS_WAVES}>" if plan.ppl > 0 else
            ("k_bs_partial_f64_grp<4, 2, 2>" if plan.ppl == -2 else "k_bs_partial_f64_grp<8, 2, 4>"))
        print(json.dumps({"workload": wl.name, "kernel": kname, "work_items": plan.nwork,
                          "launches": args.roofline_launches + 1, "avg_launch_ms": avg}), flush=True)
        return 0

    ms32, out32, plan32, partial32 = S.forward(FP32, tuned[FP32], args.steps, args.warmup, args.graph)
    rate32 = S.particle_steps / (ms32 / 1e3)
    log(f"[bench] fp32: {ms32:.1f} ms/pass, {rate32:.4g} particle-steps/s")

    fp64 = None
    err = None
    if not args.no_fp64:
        tuned[64] = S.shape(64, args.ppl, args.items_per_cu)
        ms64, out64, _, _ = S.forward(64, tuned[64], 1, 1 if args.warmup > 0 else 0, args.graph)
        rate64 = S.particle_steps / (ms64 / 1e3)
        d = np.linalg.norm(out32[:, :3] - out64[:, :3], axis=1)
        err = {"max": float(d.max() * 1e6), "rms": float(np.sqrt(np.mean(d ** 2)) * 1e6),
               "p99": float(np.percentile(d, 99) * 1e6), "median": float(np.median(d) * 1e6),
               "within_1um": float(np.mean(d <= 1e-6))}
        fp64 = {"particle_steps_per_s": rate64, "pair_evals_per_s": S.pairs_per_pass / (ms64 / 1e3),
                "ms_per_step": ms64, "ppl": tuned[64]["ppl"]}
        log(f"[bench] fp64: {ms64:.1f} ms/pass, {rate64:.4g} particle-steps/s; fp32 position error "
            f"max {err['max']:.3g} um, rms {err['rms']:.3g} um")

    # live timing of the dominant kernel (split-K Biot-Savart, FP32) on the stream it runs on, at
    # states recorded along the FP32 track (its cost depends on how close particles get to windings)
    every = max(1, S.max_steps // 10)
    rec = torch.empty((S.max_steps // every + 1, 6, S.P), dtype=torch.float64, device=engine.device())
    S.st.copy_(S.s0)
    scratch32, _ = model.workspace(plan32)
    model.track_tensors(plan32, FP32, S.st, S.kq, S.dt, S.ns_t, S.max_steps, scratch32, partial32, record=rec,
                        record_every=every)
    per = max(5, args.roofline_launches // rec.shape[0])
    avg_ms = float(np.mean([bs_launch_timing(plan32, rec[i, :3].contiguous(), partial32, per)
                            for i in range(rec.shape[0])]))
    flops = float(np.sum(np.asarray(S.npt) * model.design_nseg)) * FLOP_PER_PAIR
    achieved = flops / (avg_ms / 1e3) / 1e12
    kname = kernel_name(plan32.ppl)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[32], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS[32], "traffic": measured_traffic(wl.name, kname, plan32.nwork),
                "kernel": kname, "avg_launch_ms": avg_ms, "flop_per_launch": flops,
                "algorithmic_bytes_per_launch": float(model.nseg_padded * 48 + plan32.npt * 24
                                                      + plan32.nwork * 3 * engine.points_per_block(plan32.ppl) * 8)}
    if roofline["traffic"] is not None:  # HBM bytes per launch (counters) at this launch duration
        roofline["traffic_GB_per_s"] = roofline["traffic"] / (avg_ms / 1e3) / 1e9
    log(f"[bench] bs kernel: {avg_ms * 1e3:.1f} us/launch, {achieved:.2f} TFLOP/s "
        f"({100 * roofline['frac']:.1f}% of FP32 vector peak)")

    # field of the FP32 path against FP64 at the same recorded states (relative error over points
    # whose field is at least 1 % of the largest, where a relative figure means something)
    plan64 = model.plan(S.npt, 2)
    _, partial64 = model.workspace(plan64)
    b32, b64 = torch.empty_like(S.s0[:3]), torch.empty_like(S.s0[:3])
    dns, bns = [], []
    for i in range(rec.shape[0]):
        pts = rec[i, :3].contiguous()
        model.field_tensors(plan32, FP32, pts, partial32, b32)
        model.field_tensors(plan64, 64, pts, partial64, b64)
        dns.append((b32 - b64).norm(dim=0))
        bns.append(b64.norm(dim=0))
    dn, bn = torch.cat(dns), torch.cat(bns)
    peak, dmax = float(bn.max()), float(dn.max())
    rel_all = dn / bn
    keep = bn >= 0.01 * peak
    rel = float(rel_all.max())
    field_err = {"max_rel": rel, "max_rel_B_above_1pct_peak": float(rel_all[keep].max()), "max_abs_T": dmax,
                 "peak_T": peak, "min_B_T": float(bn.min()), "points": int(dn.numel())}
    log(f"[bench] fp32 field vs fp64: max rel {rel:.3g} over all {dn.numel()} states, max abs {dmax:.3g} T")

    extras = None
    if not args.no_extras:
        extras = {}
        P_ = Setup(W.by_name("population"), align_cct=True)
        sh = P_.shape(FP32, 0, 0)
        pm32, pout32, _, _ = P_.forward(FP32, sh, 2, 1, True)
        sh64 = P_.shape(64, 0, 0)
        pm64, pout64, _, _ = P_.forward(64, sh64, 1, 1, True)
        d = np.linalg.norm(pout32[:, :3] - pout64[:, :3], axis=1)
        fa = P_.forward_adjoint(sh, 1, 1)
        fa0 = P_.forward_adjoint(sh, 1, 1, ckpt_every=0)
        extras["population"] = {
            "designs": len(P_.bls), "segments": int(P_.model.nseg), "protons": P_.P, "rk4_steps_max": P_.max_steps,
            "fp32_particle_steps_per_s": P_.particle_steps / (pm32 / 1e3), "fp32_ms_per_pass": pm32,
            "fp32_shape": sh, "fp64_particle_steps_per_s": P_.particle_steps / (pm64 / 1e3), "fp64_ms_per_pass": pm64,
            "fp32_vs_fp64": pm64 / pm32, "position_error_um_vs_fp64": {"max": float(d.max() * 1e6),
                                                                       "rms": float(np.sqrt(np.mean(d ** 2)) * 1e6)},
            "hip_graph": True, "forward_adjoint": fa, "forward_adjoint_full_tape": fa0,
            "forward_adjoint_vs_fp64_forward": fa["particle_steps_per_s"] / (P_.particle_steps / (pm64 / 1e3)),
            "forward_adjoint_full_tape_vs_fp64_forward":
                fa0["particle_steps_per_s"] / (P_.particle_steps / (pm64 / 1e3))}
        log(f"[bench] population: fp32 {pm32:.1f} ms, fp64 {pm64:.1f} ms, fwd+adj {fa['ms_per_pass']:.1f} ms, "
            f"error max {d.max() * 1e6:.3g} um")
        del P_
        A_ = Setup(W.by_name("adjoint_section"), align_cct=True)
        sh = A_.shape(FP32, 0, 0)
        fa = A_.forward_adjoint(sh, 1, 1)
        fa0 = A_.forward_adjoint(sh, 1, 1, ckpt_every=0)
        A_P, A_seg, A_steps = A_.P, int(A_.model.nseg), A_.max_steps
        del A_
        extras["single_cct_tracking"] = tracking_leg("single_cct", passes=5)
        if cpu_cfg1 is not None:
            extras["single_cct_tracking"]["cpu_full_run"] = cpu_cfg1
            extras["single_cct_tracking"]["fp32_vs_cpu_numpy"] = \
                extras["single_cct_tracking"]["fp32"]["particle_steps_per_s"] / cpu_cfg1["value"]
        extras["hust_total"] = tracking_leg("hust_total", max_steps=2000, passes=1, ppl=2, items_per_cu=4,
                                            adjoint_steps=400)
        extras["single_cct_field_along_arc"] = field_along_arc()
        log(f"[bench] single_cct field at 1000 arc points: fp64 {extras['single_cct_field_along_arc']['fp64']['ms']:.3f} ms, "
            f"fp32 {extras['single_cct_field_along_arc']['fp32']['ms']:.3f} ms")
        gc = gradient_check()
        extras["adjoint_section"] = dict(protons=A_P, segments=A_seg, rk4_steps=A_steps, **fa,
                                         full_tape=fa0, gradient_check=gc)
        log(f"[bench] adjoint_section: fwd+adj {fa['ms_per_pass']:.1f} ms ({fa['forward_ms']:.1f} forward); "
            f"gradient vs FD max rel {gc['max_rel_err_vs_fd']:.3g}")


    out = {
        "metric": METRIC,
        "value": rate32,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "dtype_detail": "f32 Biot-Savart pair terms and 32-segment tile sums; f64 terms for tiles a point lies "
                        "among, f64 tile origins, sums across tiles and RK4 state",
        "data": "synthetic",
        "config": {"workload": wl.name, "designs": len(S.bls), "segments": int(model.nseg), "protons": S.P,
                   "rk4_steps": S.max_steps, "steps_capped": args.max_steps > 0,
                   "footstep_mm": wl.footstep * 1e3, "length_m": wl.length,
                   "parallelism": "single GPU", "ppl": plan32.ppl, "work_items": plan32.nwork, "hip_graph": args.graph,
                   "target_items": tuned[FP32]["target_items"]},
        "pair_evals_per_s": S.pairs_per_pass / (ms32 / 1e3),
        "fp64_plain": fp64,
        "speedup_vs_fp64": (rate32 / fp64["particle_steps_per_s"]) if fp64 else None,
        "position_error_um_vs_fp64": err,
        "field_error_vs_fp64": field_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
