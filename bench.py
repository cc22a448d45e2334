#!/usr/bin/env python3
"""Throughput of the MPM substep on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): 4,194,304
Drucker-Prager snow particles (E=1.4e5, nu=0.2, rho0=400, M=tan 30deg,
cohesion 500) in an 8-cell layer over a seeded six-mode sine terrain given as
a dense SDF, 256^3 grid on the unit box, APIC, gravity -9.8, dt=5e-5.  A step
is one full substep (bin, P2G, grid update with SDF BC, G2P and the FP64
fallback's fix-up) over all particles, inputs resident in HBM.  The timed
window is substeps 20..500 of the scene's 500-substep run (--warmup 20
--steps 480).  `value` = particles x steps / seconds.

Also printed in the JSON line:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average HIP-event duration on the engine's stream (the same
                window replayed on a fresh engine), against the 8 TB/s HBM
                peak (bytes per particle / per active node as in the substep
                model B = 288 N + 68 G_act: bin 20 N + 8 G, P2G 92 N + 16 G,
                grid 28 G, G2P 176 N + 16 G); valu_frac: its measured VALU
                instructions per launch against the VALU issue peak; bound:
                whichever fraction is larger
  svd_fp64_fallback  the window's lanes taking the FP64 fallback and its time
                with the fallback off
  adjoint       the reverse-mode leg of the metric on the differentiable
                batch (64 scenes of 32,768 particles on 64^3 grids, an elastic
                layer on a Drucker-Prager layer, hidden E, nu, friction angle
                per scene, proper polar SVD): forward-only time of 400
                substeps against a recorded forward of 400 substeps with a
                checkpoint every 20 plus the backward (dL/d initial state,
                dL/d material parameters)
  adjoint_single_scene  the same ratio on this scene (proper polar SVD),
                K substeps, all states kept and sqrt(K)-substep segments
  batch         the 256-scene mix (particle counts lognormal(ln 16000, 0.8)
                clipped to [2000, 64000], grids 32..96 cells, elastic / Von
                Mises / Drucker-Prager falling boxes, seed 4): the plain
                per-scene loop (one engine per scene, each stepped in turn)
                against one batch engine stepping all scenes together, eager,
                after the autotuner, and as hipGraph replays
  cpu_baseline  the FP32 CPU oracle (a restatement of the reference's shaders,
                OpenMP over the host cores) timed on a bounded sample: a few
                substeps of the same scene
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s at 4M DP particles (1 GPU); fwd+adjoint / fwd time"
HBM_PEAK_GBS = 8000.0
# VALU issue: 1024 SIMDs at 2.4 GHz, 4 cycles per wave64 instruction (16
# lanes a cycle for FP32 without packed math, and for FP64)
SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4
# HBM bytes per launch of the substep kernels over the default window, measured
# with rocprofv3 counter passes of this build (FETCH_SIZE, then WRITE_SIZE),
# scaled by what the same counters read for a 256 MiB copy kernel in the same
# process (fetch x 2.0, write x 1.0) and checked against the TCC request
# counters.  A counter pass cannot run inside the timed bench; the commands and
# the per-kernel table are in the file named below.
TRAFFIC_SOURCE = "profiles/hbm_traffic_snow_terrain_256.md"
MEASURED_TRAFFIC = {"g2p": 1495.2e6, "p2g": 613.6e6, "grid_update": 120.9e6}
# VALU instructions (wave64) per launch over the default window, from a
# rocprofv3 SQ_INSTS_VALU pass of this build (file below): with the live
# average launch time they give the fraction of the VALU issue peak.
VALU_SOURCE = "profiles/counters_snow_terrain_256.md"
MEASURED_VALU = {"g2p": 1.823e8, "p2g": 5.299e7, "grid_update": 5.518e6}
# sha256 (first 16 hex digits) of the forward kernels' sources the two tables
# above were measured on; the bench line says whether the running build's
# sources still hash to it (if not, the counter passes are due again)
COUNTER_SOURCES = ("nebula_amd/csrc/mpm_kernels.hip", "nebula_amd/csrc/nb_device_math.hpp",
                   "nebula_amd/csrc/nb_kernel_util.hpp")
COUNTER_SOURCES_SHA = "01205c9029ebebae"


def counter_sources_sha():
    import hashlib

    h = hashlib.sha256()
    for f in COUNTER_SOURCES:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, n, g_act):
    """Algorithmic bytes per launch.  g2p_grid is the grid update fused into
    G2P (every substep of a call but the last): it reads the partial sums (16 B
    per active node) instead of the updated grid and writes no grid."""
    return {
        "p2g": 92 * n + 16 * g_act,
        "g2p": 176 * n + 16 * g_act,
        "grid_update": 28 * g_act,
        "g2p_grid": 176 * n + 16 * g_act,
    }.get(name)


def adjoint_leg(sc, k):
    """Forward+adjoint over forward time for k substeps of the scene (proper
    polar SVD), with every state kept and with sqrt(k)-substep segments."""
    import dataclasses
    import math

    import numpy as np

    from nebula_amd.scenes import make_engine

    sc = dataclasses.replace(sc, corrected_svd=True)  # the caller's scene keeps its convention
    eng = make_engine(sc)
    n = sc.n
    eng.substep(sc.dt, 4)
    eng.sync()
    start = eng.getState()  # every timing starts here: the same k substeps each time
    t0 = time.perf_counter()
    eng.substep(sc.dt, k)
    eng.sync()
    t_fwd = time.perf_counter() - t0
    seed = np.random.default_rng(0).standard_normal((8, n, 4), dtype=np.float32)
    res = {"substeps": k, "fwd_ms_per_substep": round(t_fwd / k * 1e3, 4)}
    for tag, seg in (("all_states", k), ("checkpointed", max(1, math.isqrt(k)))):
        eng.setState(start)
        eng.adjointBegin(k, seg)
        t0 = time.perf_counter()
        eng.forwardRecorded(sc.dt)
        t_rec = time.perf_counter() - t0
        eng.setAdjoint(seed)
        t0 = time.perf_counter()
        eng.backward()
        t_bwd = time.perf_counter() - t0
        eng.adjointEnd()
        res[tag] = {"segment": seg, "recorded_fwd_ms_per_substep": round(t_rec / k * 1e3, 4),
                    "bwd_ms_per_substep": round(t_bwd / k * 1e3, 4), "ratio": round((t_rec + t_bwd) / t_fwd, 3)}
        log(f"[bench] adjoint ({tag}, segment {seg}): fwd {t_fwd / k * 1e3:.3f} ms, recorded fwd "
            f"{t_rec / k * 1e3:.3f} ms, bwd {t_bwd / k * 1e3:.3f} ms per substep; ratio {res[tag]['ratio']:.3f}")
    res["ratio"] = res["all_states"]["ratio"]
    return res


def host_transfer_leg(eng, sc, frame_substeps=20, reps=3):
    """Copies across the boundary for callers that hand over host arrays:
    the whole state (eight float4 arrays, 128 B per particle) each way and
    the positions alone back, through nb_mpm_set_field / nb_mpm_get_field
    from pageable numpy arrays; then whole frames of frame_substeps substeps
    (the reference's default numSubsteps) that upload the state first and
    read the positions back after, timed end to end (median of reps)."""
    n = sc.n
    state = eng.getState()
    eng.setState(state)
    eng.sync()

    def timed(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            eng.sync()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    t_down = timed(eng.getState)
    t_up = timed(lambda: eng.setState(state))
    t_pos = timed(lambda: eng.getField(0))

    def frame():
        eng.setState(state)
        eng.substep(sc.dt, frame_substeps)
        eng.getField(0)

    t_frame = timed(frame)
    gb = n * 128 / 1e9
    log(f"[bench] host copies: state up {t_up * 1e3:.1f} ms, down {t_down * 1e3:.1f} ms, positions down "
        f"{t_pos * 1e3:.1f} ms; frame of {frame_substeps} substeps with upload and readback {t_frame * 1e3:.1f} ms")
    return {"state_bytes": n * 128, "state_upload_ms": round(t_up * 1e3, 3),
            "state_upload_GBps": round(gb / t_up, 2), "state_download_ms": round(t_down * 1e3, 3),
            "state_download_GBps": round(gb / t_down, 2), "positions_download_ms": round(t_pos * 1e3, 3),
            "frame_substeps": frame_substeps, "frame_ms": round(t_frame * 1e3, 3),
            "frame_particle_substeps_per_s": round(n * frame_substeps / t_frame, 1),
            "note": "host arrays are pageable numpy memory; value is measured with the state resident in HBM"}


def window_run(sc, warmup, steps, fallback=True, stats=False, profile=False):
    """A fresh engine on the scene's initial state: warmup substeps, then the
    window of `steps` substeps, timed on the host clock with the stream
    synchronised on both sides, or with HIP events around every kernel
    (profile).  Returns (ms per substep, {kernel: ms} or None, {kernel:
    launches} or None, fallback counters or None)."""
    from nebula_amd.scenes import make_engine

    eng = make_engine(sc)
    eng.setSvdFallback(fallback, count_stats=stats)
    eng.substep(sc.dt, warmup)
    eng.sync()
    if stats:
        eng.svdFallbackStats(reset=True)
    prof = calls = None
    if profile:
        prof = eng.profileSubsteps(sc.dt, steps)
        calls = eng.profileCalls()
        ms = sum(prof.values()) / steps
    else:
        t0 = time.perf_counter()
        eng.substep(sc.dt, steps)
        eng.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
    st = eng.svdFallbackStats() if stats else None
    del eng
    return ms, prof, calls, st


def svd_fallback_leg(sc, warmup, steps, ms_on):
    """The FP64 SVD fallback of near-degenerate F (on by default, part of
    value) over the timed window: the fraction of Drucker-Prager lanes and of
    wavefronts that take it, and the window's time without it against the
    time with it (ms_on, the window's own)."""
    _, _, _, st = window_run(sc, warmup, steps, fallback=True, stats=True)
    ms_off = window_run(sc, warmup, steps, fallback=False)[0]
    from nebula_amd.mpm import SvdFallback  # noqa: F401  (the defaults below are the library's)
    lanes = max(st["dp_lanes"], 1)
    res = {"gap_rel": 1e-4, "zero_rel": 1e-6, "rounding_tol": 1e-5, "substeps": steps, "after": warmup,
           "return_map_lane_frac": round(st["fp64_return_map_lanes"] / lanes, 6),
           "stress_lane_frac": round(st["fp64_stress_lanes"] / lanes, 6),
           "wave_frac": round(st["fp64_waves"] / max(st["dp_waves"], 1), 6),
           "ms_per_substep_on": round(ms_on, 5), "ms_per_substep_off": round(ms_off, 5),
           "cost_frac": round(ms_on / ms_off - 1.0, 5)}
    log(f"[bench] FP64 SVD fallback: {res}")
    return res


SWEEP_SIZES = [(128, 8), (256, 4), (256, 8), (256, 16), (512, 8), (512, 16)]


def size_sweep_leg(k, warmup):
    """The headline scene at other sizes on the one GPU: grid resolution res
    (h = 1/res) and snow layer thickness (cells), 8 particles per cell, from
    1.05 M to 33.5 M particles; dt scales with h (5e-5 at res 256)."""
    from nebula_amd.scenes import make_engine, snow_terrain_256

    out = []
    for res, layer in SWEEP_SIZES:
        sc = snow_terrain_256(res=res, layer=layer)
        dt = sc.dt * 256.0 / res
        eng = make_engine(sc)
        eng.substep(dt, warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.substep(dt, k)
        eng.sync()
        el = time.perf_counter() - t0
        del eng
        row = {"grid": res, "layer_cells": layer, "particles": sc.n, "dt": dt, "substeps": k,
               "ms_per_substep": round(el / k * 1e3, 4), "particle_substeps_per_s": round(sc.n * k / el, 1)}
        log(f"[bench] size sweep: {res}^3 grid, {sc.n} particles: {row['ms_per_substep']:.3f} ms/substep, "
            f"{row['particle_substeps_per_s']:.4e} particle-substeps/s")
        out.append(row)
    return out


def vm_leg(sc, k, warmup):
    """The scene's Von Mises variant (q_max = 500 in place of Drucker-Prager,
    same particles, SDF and dt): a secondary line of the headline workload."""
    import dataclasses

    from nebula_amd.mpm import MaterialModel
    from nebula_amd.scenes import _material, make_engine

    vm = dataclasses.replace(sc, name="snow_terrain_256_vm",
                             materials=[_material(1.4e5, 0.2, 400.0, MaterialModel.VON_MISES, qmax=500.0)])
    log(f"[bench] Von Mises variant: {k} substeps after {warmup}")
    eng = make_engine(vm)
    eng.substep(vm.dt, warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.substep(vm.dt, k)
    eng.sync()
    el = time.perf_counter() - t0
    log(f"[bench] Von Mises variant: {el / k * 1e3:.4f} ms/substep")
    return {"workload": "snow_terrain_256 with Von Mises snow (q_max 500): 4,194,304 particles, 256^3 grid, APIC, "
                        "dt 5e-5", "substeps": k, "ms_per_substep": round(el / k * 1e3, 5),
            "particle_substeps_per_s": round(vm.n * k / el, 1)}


def diff_batch_leg(k, seg):
    """Forward+adjoint over forward time on the 64-scene differentiable batch."""
    import numpy as np

    from nebula_amd.scenes import batch_layers_64, make_batch_engine

    scenes = batch_layers_64()
    dt = scenes[0].dt
    n = sum(s.n for s in scenes)
    log(f"[bench] differentiable batch: {len(scenes)} scenes, {n} particles, {k} substeps, checkpoint every {seg}")
    eng = make_batch_engine(scenes)
    init = eng.getState()
    eng.substep(dt, 2)
    eng.sync()
    eng.setState(init)
    t0 = time.perf_counter()
    eng.substep(dt, k)
    eng.sync()
    t_fwd = time.perf_counter() - t0

    def run(tapes):
        eng.setState(init)
        eng.adjointBegin(k, seg, tapes)
        n_tapes = eng.adjointTapes()
        t0 = time.perf_counter()
        eng.forwardRecorded(dt)
        t_rec = time.perf_counter() - t0
        final = eng.getState()
        target = final.copy()
        target[0, :, :3] += np.float32(1e-3)  # loss = MSE of final positions against shifted targets
        g = np.zeros_like(final)
        g[0, :, :3] = 2.0 * (final[0, :, :3] - target[0, :, :3]) / n
        eng.setAdjoint(g)
        t0 = time.perf_counter()
        dpar = eng.backward()
        t_bwd = time.perf_counter() - t0
        grad = eng.getAdjoint()
        eng.adjointEnd()
        ratio = (t_rec + t_bwd) / t_fwd
        log(f"[bench] differentiable batch, {n_tapes} tape(s): fwd {t_fwd / k * 1e3:.3f} ms, recorded fwd "
            f"{t_rec / k * 1e3:.3f} ms, bwd {t_bwd / k * 1e3:.3f} ms per substep; ratio {ratio:.3f}")
        return n_tapes, t_rec, t_bwd, ratio, dpar, grad

    # one tape: checkpoints every `seg` substeps, every segment but the last
    # recomputed; then as many segments' tapes as fit in free HBM (the
    # gradients must be the same bits)
    _, t_rec, t_bwd, ratio, dpar, grad = run(1)
    n_res, r_rec, r_bwd, r_ratio, r_dpar, r_grad = run(0)
    same = bool(np.array_equal(grad, r_grad) and np.array_equal(np.asarray(dpar), np.asarray(r_dpar)))
    return {"workload": "64 scenes x 32,768 particles on 64^3 grids, elastic layer on Drucker-Prager layer, "
                        "hidden (E, nu, friction angle) per scene (seed 5), proper polar SVD, APIC, dt 1e-4",
            "particles": n, "substeps": k, "segment": seg, "fwd_ms_per_substep": round(t_fwd / k * 1e3, 4),
            "recorded_fwd_ms_per_substep": round(t_rec / k * 1e3, 4), "bwd_ms_per_substep": round(t_bwd / k * 1e3, 4),
            "ratio": round(ratio, 3), "finite_gradients": bool(np.isfinite(dpar).all()),
            "resident_tapes": {"tapes": n_res, "segments": -(-k // seg),
                               "recorded_fwd_ms_per_substep": round(r_rec / k * 1e3, 4),
                               "bwd_ms_per_substep": round(r_bwd / k * 1e3, 4), "ratio": round(r_ratio, 3),
                               "same_gradients": same}}


def batch_leg(k):
    """The 256-scene mix: plain per-scene loop against one batch engine."""
    from nebula_amd.scenes import batch_mix_256, make_batch_engine, make_engine

    scenes = batch_mix_256()
    dt = scenes[0].dt
    n = sum(s.n for s in scenes)
    log(f"[bench] 256-scene mix: {len(scenes)} scenes, {n} particles; plain loop, {k} substeps")
    engines = [make_engine(s) for s in scenes]
    for e in engines:
        e.substep(dt, 3)
    for e in engines:
        e.sync()
    t0 = time.perf_counter()
    for e in engines:
        e.substep(dt, k)
    for e in engines:
        e.sync()
    t_loop = (time.perf_counter() - t0) / k
    del engines
    b = make_batch_engine(scenes)
    b.substep(dt, 2)
    b.sync()
    start = b.getState()

    # every form is timed over the same substeps as the loop (3 .. 3 + k): from
    # the state after 2 substeps, one warm substep, then k timed
    def timed(step):
        b.setState(start)
        b.substep(dt, 1)
        b.sync()
        t0 = time.perf_counter()
        step()
        b.sync()
        return (time.perf_counter() - t0) / k

    t_eager = timed(lambda: b.substep(dt, k))
    tune = b.autotune(dt, trial_substeps=4)
    t_tuned = timed(lambda: b.substep(dt, k))
    b.graphSubstep(dt, 2)  # capture
    b.sync()
    t_graph = timed(lambda: b.graphSubstep(dt, k))
    best = min(t_eager, t_tuned, t_graph)
    log(f"[bench] 256-scene mix per substep: loop {t_loop * 1e3:.3f} ms, batch {t_eager * 1e3:.3f} ms, "
        f"tuned {t_tuned * 1e3:.3f} ms, graph {t_graph * 1e3:.3f} ms; speedup {t_loop / best:.2f}x ({tune})")
    return {"workload": "256 falling-box scenes, particle counts lognormal(ln 16000, 0.8) clipped to [2000, 64000], "
                        "grids 32..96 cells at h = 1/64, elastic / Von Mises / Drucker-Prager, seed 4, APIC, dt 1e-4",
            "scenes": len(scenes), "particles": n, "substeps": k,
            "loop_ms_per_substep": round(t_loop * 1e3, 4), "batch_ms_per_substep": round(t_eager * 1e3, 4),
            "tuned_ms_per_substep": round(t_tuned * 1e3, 4), "graph_ms_per_substep": round(t_graph * 1e3, 4),
            "particle_substeps_per_s": round(n / best, 1), "speedup_vs_loop": round(t_loop / best, 3),
            "tuning": {"tile_blocks": tune.tile_blocks, "g2p_block": tune.g2p_block, "p2g_item": tune.p2g_item}}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=5, help="substeps of the CPU oracle sample (0 = skip)")
    ap.add_argument("--adjoint-substeps", type=int, default=16,
                    help="substeps of the single-scene reverse-mode leg (0 = skip)")
    ap.add_argument("--diff-substeps", type=int, default=400,
                    help="substeps of the differentiable-batch reverse-mode leg (0 = skip)")
    ap.add_argument("--diff-segment", type=int, default=20, help="checkpoint spacing of that leg")
    ap.add_argument("--batch-substeps", type=int, default=20, help="substeps of the 256-scene leg (0 = skip)")
    ap.add_argument("--vm-steps", type=int, default=100, help="substeps of the Von Mises variant (0 = skip)")
    ap.add_argument("--fallback", type=int, default=1,
                    help="count the FP64 SVD fallback's lanes and time the window without it (0 = skip)")
    ap.add_argument("--host-frames", type=int, default=1, help="time host copies and whole frames with them (0 = skip)")
    ap.add_argument("--sweep-substeps", type=int, default=0,
                    help="substeps per size of the scene-size sweep, 1.05 M to 33.5 M particles (0 = skip)")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="also run a plain copy kernel of this many MiB (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the single-scene substep is not sharded across GPUs)", file=sys.stderr)
        sys.exit(2)

    from nebula_amd import mpm
    from nebula_amd.scenes import make_engine, snow_terrain_256

    log("[bench] building scene snow_terrain_256")
    sc = snow_terrain_256()
    n = sc.n
    eng = make_engine(sc)
    log(f"[bench] {n} particles, grid {sc.grid_res()}, warmup {args.warmup}")
    eng.substep(sc.dt, args.warmup)
    eng.sync()
    log(f"[bench] timing {args.steps} substeps")
    t0 = time.perf_counter()
    eng.substep(sc.dt, args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    log(f"[bench] {ms_per_step:.4f} ms/substep, {value:.4e} particle-substeps/s")

    g_act = eng.activeNodeCount()
    tiles = eng.activeTileCount()
    # the same window again on a fresh engine, every kernel timed with HIP
    # events on the engine's stream
    log(f"[bench] profiling the same {args.steps} substeps per kernel")
    _, prof, calls, _ = window_run(sc, args.warmup, args.steps, profile=True)
    avg_ms = {k: v / max(calls.get(k, 0), 1) for k, v in prof.items() if calls.get(k, 0)}
    per_substep_ms = {k: v / args.steps for k, v in prof.items()}
    dom = max(("p2g", "g2p", "grid_update", "g2p_grid"), key=lambda k: avg_ms.get(k, 0.0))
    dbytes = kernel_bytes(dom, n, g_act)
    achieved = dbytes / (avg_ms[dom] * 1e-3) / 1e9
    b_alg = 288 * n + 68 * g_act
    valu = MEASURED_VALU.get(dom)
    valu_frac = valu * VALU_CYCLES / (SIMDS * CLOCK_HZ * avg_ms[dom] * 1e-3) if valu else None
    hbm_frac = achieved / HBM_PEAK_GBS
    roofline = {"kernel": dom, "bound": "valu" if valu_frac and valu_frac > hbm_frac else "hbm",
                "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_frac, 4), "traffic": MEASURED_TRAFFIC.get(dom),
                "traffic_source": TRAFFIC_SOURCE if dom in MEASURED_TRAFFIC else None,
                "valu_frac": round(valu_frac, 4) if valu_frac else None,
                "valu_insts_per_launch": valu, "valu_source": VALU_SOURCE if valu else None,
                "bytes_per_launch": dbytes, "avg_launch_ms": round(avg_ms[dom], 5),
                "counters_measured_on_sources": COUNTER_SOURCES_SHA,
                "counters_match_this_build": counter_sources_sha() == COUNTER_SOURCES_SHA}
    substep_rate = b_alg / (ms_per_step * 1e-3) / 1e9

    if args.calibration_copy_mb > 0:
        eng.calibrationCopy(args.calibration_copy_mb << 20, 3)

    fallback = svd_fallback_leg(sc, args.warmup, args.steps, ms_per_step) if args.fallback else None
    host = host_transfer_leg(eng, sc) if args.host_frames else None
    adjoint1 = adjoint_leg(sc, args.adjoint_substeps) if args.adjoint_substeps > 0 else None
    del eng
    vm = vm_leg(sc, args.vm_steps, args.warmup) if args.vm_steps > 0 else None
    adjoint = diff_batch_leg(args.diff_substeps, args.diff_segment) if args.diff_substeps > 0 else None
    batch = batch_leg(args.batch_substeps) if args.batch_substeps > 0 else None
    sweep = size_sweep_leg(args.sweep_substeps, args.warmup) if args.sweep_substeps > 0 else None

    cpu = None
    if args.cpu_steps > 0:
        import oracle

        if not os.path.exists(oracle.LIB_PATH):
            oracle.build()
        log(f"[bench] CPU oracle sample: {args.cpu_steps} substeps of the same scene")
        scn = oracle.make_scene(sc.dt, sc.cfg.cellSize, sc.grid_res(), (0.0, 0.0, 0.0), tuple(sc.cfg.domainSize),
                                sc.flip_ratio)
        sim = oracle.Sim(scn, mpm.materials_bytes(sc.materials), sc.colliders.to_bytes(),
                         mpm.forces_bytes(sc.forces), sc.sdf, 32)
        sim.set_state(sc.fields)
        secs = sim.substeps(args.cpu_steps)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(n * args.cpu_steps / secs, 1), "unit": "particle-substeps/s", "cores": cores,
               "kind": "port",
               "sample": f"FP32 CPU oracle (C++/OpenMP restatement of the reference shaders), "
                         f"{args.cpu_steps} substeps of the full {n}-particle scene, {secs:.2f} s"}
        log(f"[bench] cpu {cpu['value']:.4e} particle-substeps/s on {cores} threads")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: seeded six-mode sine terrain SDF (seed 3), 8 jittered particles per cell",
        "config": {"workload": "snow_terrain_256: 4,194,304 Drucker-Prager snow particles over a terrain SDF, "
                               "256^3 grid, APIC, dt 5e-5", "particles": n, "grid": list(sc.grid_res()),
                   "active_nodes": g_act, "active_tiles": tiles, "parallelism": "single GPU"},
        "roofline": roofline,
        "substep_roofline": {"bytes_per_step": b_alg, "achieved": round(substep_rate, 2), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(substep_rate / HBM_PEAK_GBS, 4)},
        "kernels_avg_ms": {k: round(v, 5) for k, v in avg_ms.items()},
        "kernels_ms_per_substep": {k: round(v, 5) for k, v in per_substep_ms.items() if v > 0.0},
        "kernel_launches": {k: v for k, v in calls.items() if v},
        "svd_fp64_fallback": fallback,
        "host_transfer": host,
        "von_mises_variant": vm,
        "adjoint": adjoint,
        "adjoint_single_scene": adjoint1,
        "batch": batch,
        "size_sweep": sweep,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
