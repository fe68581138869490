"""Timing of the reference CPU path as the oracle restates it (test infrastructure only; run
by bench.py's cpu_baseline leg in a child process that never touches the GPU).

The numpy restatement follows the reference op for op: per-particle sequential RK4
(final_code/packages/base_utils.py:263-311) with the ParticleRunner callback
(particles.py:514-537) and the numpy Biot-Savart sum of every CCT in its local frame
(cct.py:249-285), one process per particle as run_only(concurrency_level=os.cpu_count())
does (particles.py:617-632).  It is timed in one process and in a process pool; the C
restatement (OpenMP over particles) is timed beside it as a second, labelled point.

    python -m oracle.cpu_baseline WORKLOAD N1 STEPS1 NPOOL STEPSPOOL WORKERS
prints one JSON object.
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys
import time

import numpy

from . import c_oracle, cct_oracle as O, workload_builder as WB

_BL = None


def _workloads():
    """cctpy_amd/workloads.py loaded from its file (plain data: the builder operations and the
    seeded beams), so that this process loads no GPU library."""
    import importlib.util
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cctpy_amd", "workloads.py")
    spec = importlib.util.spec_from_file_location("cctpy_amd_workloads_data", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod  # its dataclasses look their module up
    spec.loader.exec_module(mod)
    return mod


def _setup(workload: str):
    wl = _workloads().by_name(workload)
    bl = WB.build_beamline(wl.designs[0])
    b = wl.beams[0]
    ip = O.create_proton_along(bl.traj, b.s0, b.kinetic_mev)
    fr = ip.natural_frame()
    ps = [O.from_phase_space(ip, fr, c[0], c[1], c[2], c[3], 0.0, c[4]) for c in b.coords]
    return wl, bl, ps


def _init(workload: str):
    global _BL
    _BL = _setup(workload)


def _run_one(args):
    i, steps = args
    wl, bl, ps = _BL
    p = ps[i].copy()
    O.run_only(p, bl.field, steps * wl.footstep, wl.footstep)
    return p.position


def main(argv) -> int:
    workload, n1, s1, npool, spool, workers = argv[0], int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]), \
        int(argv[5])
    wl, bl, ps = _setup(workload)
    nseg = sum(int(m.kl.shape[0]) for m in bl.magnets if isinstance(m, O.CCTMagnet))
    # one process: n1 protons one after the other, s1 RK4 steps each (length = s1 footsteps)
    t0 = time.perf_counter()
    for i in range(n1):
        O.run_only(ps[i].copy(), bl.field, s1 * wl.footstep, wl.footstep)
    el1 = time.perf_counter() - t0
    # process pool: one task per proton, npool protons of spool steps
    ctx = mp.get_context("fork")
    with ctx.Pool(workers, initializer=_init, initargs=(workload,)) as pool:
        pool.map(_run_one, [(0, 1)] * workers)  # every worker built its beamline
        t0 = time.perf_counter()
        pool.map(_run_one, [(i, spool) for i in range(npool)], chunksize=1)
        elp = time.perf_counter() - t0
    # the C restatement (OpenMP over protons) on the pool's sample
    mags = O.beamline_arrays(bl.magnets)
    st = numpy.array([list(p.position) + list(p.velocity) for p in ps[:npool]])
    kq = numpy.array([p.e / p.rm for p in ps[:npool]])
    dts = numpy.array([O.rk4_steps(spool * wl.footstep, wl.footstep, p.speed)[1] for p in ps[:npool]])
    ns = numpy.array([O.rk4_steps(spool * wl.footstep, wl.footstep, p.speed)[0] for p in ps[:npool]], numpy.int64)
    c_oracle.track(mags, st[:1], kq[:1], dts[:1], numpy.ones(1, numpy.int64), threads=workers)  # warm
    t0 = time.perf_counter()
    c_oracle.track(mags, st, kq, dts, ns, threads=workers)
    elc = time.perf_counter() - t0
    print(json.dumps({
        "workload": workload, "segments": nseg, "workers": workers,
        "numpy_1proc": {"particle_steps": n1 * s1, "seconds": el1, "particle_steps_per_s": n1 * s1 / el1},
        "numpy_pool": {"particle_steps": npool * spool, "seconds": elp, "particle_steps_per_s": npool * spool / elp},
        "c_openmp": {"particle_steps": int(ns.sum()), "seconds": elc, "particle_steps_per_s": float(ns.sum()) / elc},
    }))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
