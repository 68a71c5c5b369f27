"""Benchmark: batched persistent ADAM minimization of the 262,144-pose / 64-ligand workload
(2000-bead receptor, 96^3 + 64^3 grids, ligands of 5-60 beads, 200 ADAM steps) on one GPU.

One timed step = one persistent launch that runs all 200 ADAM steps of all 262,144 poses
(the pose state is reset from HBM-resident initial poses before each step).  Prints one
JSON line.  Also measured in the same run: the plain per-job form (one ligand at a time,
one energrad + one ADAM launch per ADAM step) and the oracle's FP32 CPU twin on a
bounded sample (cpu_baseline).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pose-steps/sec batched vs plain at 1e-4 rel energy; p50/p99 dock-request latency"
HBM_PEAK_GBS = 8000.0
# VALU issue limit: each of the 4 SIMDs of a CU issues one wave64 VALU instruction per 4
# cycles (16 lanes a cycle), i.e. one per CU per cycle: 256 CUs x 2.4 GHz peak engine clock
VALU_PEAK_PER_S = 256 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _counter_per_dispatch(csv_path, counter):
    """{dispatch id: (kernel name, summed counter value)} of a rocprofv3 counter CSV."""
    import csv
    out = {}
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            k = int(r["Dispatch_Id"])
            name, v = out.get(k, (r["Kernel_Name"], 0.0))
            out[k] = (name, v + float(r["Counter_Value"]))
    return out


def scoring_rates(dev, wpb):
    """Side figures beside the minimizer: one energrad of the C2 workload (131,072 poses of a
    21-bead ligand against the C2 grid) and one pair-table adjoint of the C4 workload (65,536
    poses), each the best of three launches timed with HIP events on the launch stream."""
    import torch

    from attract_jax_amd import _lib
    from attract_jax_amd.autodiff import param_adjoint
    from attract_jax_amd.pose import mat_from_qt
    from attract_jax_amd.workloads import build_system, c2, c4

    def best_ms(fn, stream):
        fn()
        ms = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return min(ms)

    stream = torch.cuda.current_stream(dev)
    w2 = c2()
    system, ligs = build_system(w2, dev, launch=_lib.AtjLaunch(wpb, 0, 0))
    mat = torch.from_numpy(mat_from_qt(w2.quat, w2.trans).astype(np.float32)).to(dev)
    ms2 = best_ms(lambda: system.energrad(mat, ligs[0]), stream)
    w4 = c4()
    m4 = torch.from_numpy(mat_from_qt(w4.quat, w4.trans).astype(np.float32)).to(dev).reshape(-1, 16)
    wts = torch.ones(m4.shape[0], device=dev)
    ms4 = best_ms(lambda: param_adjoint(system, ligs[0], m4, system.ff, wts, system.flags), stream)
    log(f"[bench] scoring: C2 energrad {ms2:.3f} ms, C4 adjoint {ms4:.3f} ms")
    # test5's protocol (test5.py:14-22,450-458: seven stages, 375 adabelief steps with the gravity
    # restraint in the first four) over the C2 starting poses, one launch per stage
    from attract_jax_amd.minimize import BatchPlan, PoseState
    from attract_jax_amd.protocol import SCHEMA_TEST5, run_schedule
    plan = BatchPlan(system, ligs, [len(w2.quat)], 1, 1e-3, 1e-3)
    st5 = PoseState.create(w2.quat, w2.trans, dev)
    init5 = st5.pose.clone()

    def test5_run():
        st5.pose.copy_(init5)
        run_schedule(system, plan, st5, SCHEMA_TEST5)

    ms5 = best_ms(test5_run, stream)
    steps5 = int(sum(e for e, _, _ in SCHEMA_TEST5))
    log(f"[bench] test5 protocol: {len(w2.quat)} structures x {steps5} steps in {ms5:.1f} ms")
    # ligand RMSD of every C2 pose against the ligand's own frame (test-rmsd.py:13-17), through the C-ABI
    from attract_jax_amd.device import _ptr, _stream
    lib = _lib.load()
    lig_xyz = torch.as_tensor(np.asarray(w2.ligands[0].coor, np.float32), device=dev).contiguous()
    rmsd_out = torch.empty(mat.shape[0], dtype=torch.float32, device=dev)
    mat16 = mat.reshape(-1, 16).contiguous()
    ms_rmsd = best_ms(lambda: _lib.check(lib.atj_ligand_rmsd(_ptr(mat16), mat16.shape[0], _ptr(lig_xyz), _ptr(lig_xyz),
                                                             lig_xyz.shape[0], _ptr(rmsd_out), _stream(dev))), stream)
    return {"c2_energrad_poses_per_s": len(w2.quat) / (ms2 * 1e-3), "c2_energrad_ms": ms2, "c2_poses": len(w2.quat),
            "c4_adjoint_poses_per_s": len(w4.quat) / (ms4 * 1e-3), "c4_adjoint_ms": ms4, "c4_poses": len(w4.quat),
            "note": "energies + dE/dM of every pose (C2); sum over poses of dE/d(rc, ac, emin) with its partial-sum "
                    "reduction (C4); best of 3 launches",
            "test5_protocol": {"structures": len(w2.quat), "steps": steps5, "ms": ms5,
                               "structures_per_hour": len(w2.quat) / (ms5 * 1e-3) * 3600.0,
                               "pose_steps_per_s": len(w2.quat) * steps5 / (ms5 * 1e-3),
                               "workload": "the C2 system's 131,072 starting poses of its 21-bead ligand through "
                                           "test5's seven-stage adabelief + gravity schedule, one launch per "
                                           "stage; best of 3"},
            "ligand_rmsd": {"poses": int(mat16.shape[0]), "beads": int(lig_xyz.shape[0]), "ms": ms_rmsd,
                            "poses_per_s": mat16.shape[0] / (ms_rmsd * 1e-3)},
            "grid_build": grid_build_times(dev, best_ms, stream)}


def grid_build_times(dev, best_ms, stream):
    """The device grid builder on serving receptors of 500, 2000 and 4000 beads (32 types, inner
    grid at 0.9 A, outer grid at 1.8 A): potentials and neighbour lists, plain and corner-packed,
    and the lists alone; best of 3, HIP events on the build stream (the list allocation reads
    the list total back to the host once per build, which the time includes)."""
    from attract_jax_amd import _lib
    from attract_jax_amd.device import DeviceGrid, ForceField, Receptor
    from attract_jax_amd.workloads import NTYPES, c5_receptor

    out = []
    for nrec in (500, 2000, 4000):
        r = c5_receptor(("build", nrec), nrec, np.random.default_rng(nrec))
        ff = ForceField.create(r.rc, r.ac, r.ivor, device=dev)
        rec = Receptor.create(r.coor, r.types - 1, dev)
        row = {"receptor_beads": nrec, "inner_dim": r.grid_kw["dim"][0], "outer_dim": r.grid_kw["dim2"][0]}
        for name, kw in (("plain_ms", dict(layout=_lib.ATJ_LAYOUT_PLAIN)),
                         ("corner_ms", dict(layout=_lib.ATJ_LAYOUT_CORNER)),
                         ("lists_only_ms", dict(layout=_lib.ATJ_LAYOUT_PLAIN, potentials=False))):
            row[name] = best_ms(lambda: DeviceGrid.build(rec, ff, np.arange(NTYPES), device=dev, **kw, **r.grid_kw),
                                stream)
        log(f"[bench] grid build {nrec} beads, inner {row['inner_dim']}^3: plain {row['plain_ms']:.1f} ms, "
            f"corner {row['corner_ms']:.1f} ms, lists only {row['lists_only_ms']:.1f} ms")
        out.append(row)
    return out


def _child_command(args, wpb, opt):
    """A one-launch run of this workload that times nothing else (the profiled child)."""
    return [sys.executable, "-u", os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-cpu",
            "--no-plain", "--no-share", "--serve-jobs", "0", "--no-traffic", "--no-scoring", "--copy-calibration",
            "--waves",
            str(wpb), "--launch-options", str(opt), "--ligands", str(args.ligands), "--poses-per-ligand",
            str(args.poses_per_ligand), "--adam-steps", str(args.adam_steps), "--nshell", str(args.nshell)]


def _counter_pass(child, ctr):
    """One rocprofv3 --pmc pass of `child` for counter `ctr`: {dispatch id: (kernel, value)},
    or None when rocprofv3 is unavailable, times out or writes no counter file."""
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    env = dict(os.environ, TMPDIR="/tmp")
    tmp = tempfile.mkdtemp(prefix="atj_pmc_", dir="/tmp")
    log(f"[bench] counter pass {ctr}")
    proc = subprocess.Popen([exe, "--pmc", ctr, "--output-format", "csv", "-d", tmp, "-o", "c", "--", *child],
                            stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=env, cwd="/tmp",
                            start_new_session=True)
    try:
        status = proc.wait(timeout=240)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, 9)  # the profiler and the profiled run
        proc.wait()
        log(f"[bench] counter pass {ctr} timed out")
        shutil.rmtree(tmp, ignore_errors=True)
        return None
    # the profiler may exit non-zero after writing complete results; the CSV decides
    files = glob.glob(os.path.join(tmp, "**", "c_counter_collection.csv"), recursive=True)
    disp = _counter_per_dispatch(files[0], ctr) if files else None
    shutil.rmtree(tmp, ignore_errors=True)
    if not disp:
        log(f"[bench] counter pass {ctr} failed (status {status})")
    return disp or None


def _pose_dispatch(disp):
    pose = [disp[k][1] for k in sorted(disp) if "pose_kernel" in disp[k][0]]
    return pose[-1] if pose else None


def measure_traffic(args, wpb, opt):
    """HBM traffic of one minimizer launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the
    timed pose-kernel dispatch of a one-launch child run of this workload, one rocprofv3 pass
    per counter, each scaled by the value the same counter reads for a 1 GiB device copy in
    the same run (1 GiB read, 1 GiB written).  Returns None when rocprofv3 is unavailable or a
    pass fails."""
    child = _child_command(args, wpb, opt)
    res = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        disp = _counter_pass(child, ctr)
        if disp is None:
            return None
        copy = [v for name, v in disp.values() if "copyBuffer" in name]
        pose = _pose_dispatch(disp)
        if not copy or pose is None or max(copy) <= 0:
            return None
        scale = (2 ** 30 / 1024) / max(copy)  # KiB a 1 GiB copy should read (write), over what it reads
        res[ctr] = (pose * scale * 1024, scale)
    return {"fetch_bytes": res["FETCH_SIZE"][0], "write_bytes": res["WRITE_SIZE"][0],
            "fetch_scale": res["FETCH_SIZE"][1], "write_scale": res["WRITE_SIZE"][1]}


def measure_valu(args, wpb, opt):
    """VALU wave-instructions of one minimizer launch: SQ_INSTS_VALU of the pose-kernel dispatch
    of a one-launch child run of this workload (one rocprofv3 pass), or None."""
    disp = _counter_pass(_child_command(args, wpb, opt), "SQ_INSTS_VALU")
    return None if disp is None else _pose_dispatch(disp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ligands", type=int, default=64)
    ap.add_argument("--poses-per-ligand", type=int, default=4096)
    ap.add_argument("--adam-steps", type=int, default=200)
    ap.add_argument("--plain-ligands", type=int, default=64, help="ligands timed in the plain per-job form")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-share", action="store_true", help="skip one GPU's share of the workload on eight GPUs")
    ap.add_argument("--waves", type=int, default=0, help="waves per block (0 = autotuned)")
    ap.add_argument("--launch-options", type=int, default=0, help="ATJ_LAUNCH_* bits when --waves is given")
    ap.add_argument("--nshell", type=int, default=6, help="distance shells of the neighbour lists (0 = whole lists)")
    ap.add_argument("--serve-jobs", type=int, default=200, help="jobs of the serving run (0 = skip)")
    ap.add_argument("--serve-budget-gib", type=float, default=48.0,
                    help="HBM the serving loop may hold (grids + pose slots)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean job arrivals per second")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy (scale reference for FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-scoring", action="store_true",
                    help="skip the single-shot energrad (C2) and pair-table adjoint (C4) timings")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes that measure the minimizer's HBM traffic and VALU "
                         "instructions")
    ap.add_argument("--phase-clocks", action="store_true",
                    help="load the library built with ATJ_PHASE_CLOCKS, run one minimizer launch and print the "
                         "share of wavefront time in each phase (no timed launches)")
    args = ap.parse_args()
    if args.phase_clocks:
        os.environ["ATTRACT_AMD_LIB"] = os.path.join(ROOT, "attract_jax_amd", "libattract_amd_clocks.so")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (poses shard trivially across "
                                   "GPUs with a replicated grid, but no multi-GPU path is built)"}))
        sys.exit(2)

    import torch

    from attract_jax_amd import _lib
    from attract_jax_amd.autotune import autotune_minimize
    from attract_jax_amd.minimize import BatchPlan, PoseState, minimize, minimize_plain
    from attract_jax_amd.workloads import build_system, c3

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    w = c3(args.ligands, args.poses_per_ligand, args.adam_steps)
    log(f"[bench] workload generated in {time.time() - t0:.1f}s: {w.name}")
    t0 = time.time()
    system, ligs = build_system(w, dev, keep_plain=True, nshell=args.nshell)
    torch.cuda.synchronize()
    log(f"[bench] grid built on the GPU in {time.time() - t0:.2f}s ({system.grid.nbytes() / 1e9:.2f} GB resident)")
    plan = BatchPlan(system, ligs, w.pose_counts, w.nsteps, w.lr_rot, w.lr_trans)
    init = PoseState.create(w.quat, w.trans, dev)
    st = PoseState.create(w.quat, w.trans, dev)

    def reset():
        st.pose.copy_(init.pose)
        st.m.zero_()
        st.v.zero_()
        st.count.zero_()

    if args.waves:
        system.launch = _lib.AtjLaunch(args.waves, 0, args.launch_options)
    else:
        from attract_jax_amd.score import ScoringSystem
        plain = ScoringSystem(system.grid.with_layout(_lib.ATJ_LAYOUT_PLAIN), system.ff, system.receptor)
        layout, wpb, opt = autotune_minimize({_lib.ATJ_LAYOUT_CORNER: system, _lib.ATJ_LAYOUT_PLAIN: plain}, plan, w,
                                             dev, log=log)
        if layout == _lib.ATJ_LAYOUT_PLAIN:
            system = plain
        system.launch = _lib.AtjLaunch(wpb, 0, opt)
    log(f"[bench] layout {system.grid.layout}, waves per block: {system.launch.waves_per_block}, options "
        f"{system.launch.options:#x}")

    stream = torch.cuda.current_stream(dev)
    for i in range(args.warmup):
        reset()
        minimize(system, plan, st, w.nsteps)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")

    # stats of one launch (algorithmic bytes), measured outside the timed loop
    reset()
    system.stats.zero_()
    minimize(system, plan, st, w.nsteps, stats=True)
    torch.cuda.synchronize()
    stats = system.stats.cpu().numpy().astype(np.float64)
    E_final = st.energy.cpu().numpy()
    if args.phase_clocks:
        clk = stats[_lib.ATJ_NSTATS:]
        split = {name: float(c / clk.sum()) for name, c in zip(_lib.PHASES, clk)}
        print(json.dumps({"phase_split": split, "wavefront_clock_cycles": float(clk.sum()),
                          "workload": w.name, "layout": int(system.grid.layout),
                          "waves_per_block": system.launch.waves_per_block,
                          "options": system.launch.options}), flush=True)
        return
    stats = stats[:_lib.ATJ_NSTATS]

    ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        reset()
        ev0[i].record(stream)
        minimize(system, plan, st, w.nsteps)
        ev1[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = [a.elapsed_time(b) for a, b in zip(ev0, ev1)]
    log(f"[bench] timed {args.steps} launches: wall {wall:.3f}s, kernel ms {['%.1f' % k for k in kernel_ms]}")
    pose_steps = plan.N * w.nsteps
    value = pose_steps * args.steps / wall
    k_ms = float(np.mean(kernel_ms))

    grid_beads, inner_beads, listed, pairs = stats
    algo_bytes = 128.0 * grid_beads + 4.0 * inner_beads + 2.0 * listed
    achieved = algo_bytes / (k_ms * 1e-3) / 1e9
    flops = 110.0 * grid_beads + 25.0 * pairs

    result = {
        "metric": METRIC, "value": value, "unit": "pose-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "poses": plan.N, "ligands": args.ligands, "adam_steps": w.nsteps,
                   "receptor_beads": len(w.coor_rec), "grid": "inner 96^3 @0.9A + outer 64^3 @1.8A, 32 types",
                   "layout": "corner-packed" if system.grid.layout == _lib.ATJ_LAYOUT_CORNER else "plain",
                   "waves_per_block": system.launch.waves_per_block, "list_shells": system.grid.nshell,
                   "poses_per_claim": (system.launch.options >> 4) & 0xF or "auto",
                   "tail_claims": ({"poses_per_claim": (system.launch.options >> 8) & 0xF,
                                    "rounds": (system.launch.options >> 12) & 0xF or 1}
                                   if system.launch.options & _lib.ATJ_LAUNCH_TAIL_MASK else None),
                   "receptor_path": "L1" if system.launch.options & _lib.ATJ_LAUNCH_REC_GLOBAL else "LDS"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "pose_kernel<1> (persistent minimizer)", "kernel_ms": k_ms,
                     "algorithmic_bytes_per_launch": algo_bytes},
        "per_launch_counts": {"grid_bead_evals": grid_beads, "inner_bead_evals": inner_beads,
                              "listed_neighbours": listed, "evaluated_pairs": pairs,
                              "pairs_per_pose_step": pairs / pose_steps, "est_flops": flops,
                              "est_tflops": flops / (k_ms * 1e-3) / 1e12},
        "energy_finite_frac": float(np.isfinite(E_final).mean()),
    }

    if not args.no_share:
        # one GPU's share of the workload on eight GPUs (poses sharded, grid replicated): the first
        # eighth of every ligand's poses, one launch of all steps, best of three
        counts = [max(1, c // 8) for c in w.pose_counts]
        idx = np.concatenate([np.arange(plan.offsets[k], plan.offsets[k] + counts[k]) for k in range(len(counts))])
        sub = BatchPlan(system, ligs, counts, w.nsteps, w.lr_rot, w.lr_trans)
        sst = PoseState.create(w.quat[idx], w.trans[idx], dev)
        sinit = sst.pose.clone()
        # with 8 poses per wavefront, smaller claims even out ligands of different sizes: the
        # full workload's claim options and two with smaller claims, the fastest kept
        full_launch = system.launch
        C, T = _lib.ATJ_LAUNCH_CLAIM, _lib.ATJ_LAUNCH_TAIL
        keep = full_launch.options & _lib.ATJ_LAUNCH_REC_GLOBAL
        share_ms, share_opt = float("inf"), None
        for opt in (full_launch.options, keep | C(2), keep | C(8) | T(2, 2)):
            system.launch = _lib.AtjLaunch(full_launch.waves_per_block, full_launch.blocks, opt)
            ms = []
            for _ in range(4):
                sst.pose.copy_(sinit)
                sst.m.zero_()
                sst.v.zero_()
                sst.count.zero_()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                minimize(system, sub, sst, w.nsteps)
                e1.record(stream)
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
            if min(ms[1:]) < share_ms:
                share_ms, share_opt = min(ms[1:]), opt
        system.launch = full_launch
        result["one_gpu_share_of_8"] = {
            "poses": int(sub.N), "kernel_ms": share_ms, "pose_steps_per_s": sub.N * w.nsteps / (share_ms * 1e-3),
            "vs_full_rate": (sub.N * w.nsteps / (share_ms * 1e-3)) / (pose_steps / (k_ms * 1e-3)),
            "launch_options": share_opt}
        log(f"[bench] one GPU's share of eight ({sub.N} poses): {share_ms:.2f} ms, "
            f"{result['one_gpu_share_of_8']['vs_full_rate']:.2f} of the full-size rate")

    if not args.no_traffic:
        tr = measure_traffic(args, system.launch.waves_per_block, system.launch.options)
        if tr is not None:
            result["roofline"]["traffic"] = tr["fetch_bytes"] + tr["write_bytes"]
            result["roofline"]["traffic_detail"] = dict(
                tr, method="rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of the timed pose-kernel dispatch of a "
                           "one-launch run; KiB counts scaled by a 1 GiB device copy measured in the same run")
            log(f"[bench] traffic: {result['roofline']['traffic'] / 1e9:.1f} GB per launch "
                f"(algorithmic {algo_bytes / 1e9:.1f} GB)")
        nv = measure_valu(args, system.launch.waves_per_block, system.launch.options)
        if nv is not None:
            rate = nv / (k_ms * 1e-3)
            result["valu_issue"] = {
                "bound": "valu", "achieved": rate, "peak": VALU_PEAK_PER_S, "unit": "wave64 VALU instructions/s",
                "frac": rate / VALU_PEAK_PER_S, "instructions_per_launch": nv, "per_pose_step": nv / pose_steps,
                "method": "rocprofv3 --pmc SQ_INSTS_VALU of the pose-kernel dispatch of a one-launch run, over the "
                          "timed kernel_ms; peak = 256 CUs x one wave64 VALU issue per cycle x 2.4 GHz"}
            log(f"[bench] VALU: {nv / pose_steps:.0f} per pose-step, {rate / VALU_PEAK_PER_S:.2f} of the issue limit")

    if not args.no_plain:
        nl = min(args.plain_ligands, len(ligs))
        pst = 0
        states = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nl):
            sl = slice(int(plan.offsets[k]), int(plan.offsets[k + 1]))
            ps = PoseState.create(init.pose[sl, :4], init.pose[sl, 4:], dev)
            minimize_plain(system, ligs[k], ps, w.nsteps, w.lr_rot, w.lr_trans)
            pst += ps.N * w.nsteps
            states.append(ps)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain = pst / tp
        result["plain_pose_steps_per_s"] = plain
        result["plain_sample"] = f"{nl} of {len(ligs)} ligands x {w.pose_counts[0]} poses x {w.nsteps} steps"
        result["plain_launch"] = (f"{system.launch.waves_per_block} waves per block, poses per claim chosen by "
                                  f"each one-step launch (the minimizer's claim options left out)")
        result["batched_vs_plain"] = value / plain
        # same poses: batched result vs plain result, every ligand timed
        worst = 0.0
        for k, ps in enumerate(states):
            ref = st.energy[int(plan.offsets[k]):int(plan.offsets[k + 1])].cpu().numpy()
            ps_e = ps.energy.cpu().numpy()
            worst = max(worst, float(np.max(np.abs(ref - ps_e) / np.maximum(np.abs(ps_e), 1.0))))
        result["batched_vs_plain_max_rel_energy"] = worst
        log(f"[bench] plain form: {plain:.3e} pose-steps/s over {tp:.2f}s")
        # the same per-job launches captured in one HIP graph per ligand (captured before the
        # timed region), then all graphs replayed: the launches without their host work
        from attract_jax_amd.minimize import PlainGraph
        graphs = []
        for k in range(nl):
            sl = slice(int(plan.offsets[k]), int(plan.offsets[k + 1]))
            ps = PoseState.create(init.pose[sl, :4], init.pose[sl, 4:], dev)
            graphs.append(PlainGraph(system, ligs[k], ps, w.nsteps, w.lr_rot, w.lr_trans))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for g in graphs:
            g.replay()
        torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        same = all(torch.equal(g.state.energy, ps.energy) and torch.equal(g.state.pose, ps.pose)
                   for g, ps in zip(graphs, states))
        result["plain_graph_pose_steps_per_s"] = pst / tg
        result["plain_graph_note"] = ("the plain form's launches captured once per ligand in a HIP graph "
                                      f"({3 * w.nsteps} kernel nodes: per step a counter reset, an energrad and an optimizer launch) "
                                      "and replayed; capture outside the timed region")
        result["plain_graph_equals_plain"] = bool(same)
        result["batched_vs_plain_graph"] = value / (pst / tg)
        log(f"[bench] plain form in HIP graphs: {pst / tg:.3e} pose-steps/s over {tg:.2f}s, equal to eager: {same}")
        del graphs

    if not args.no_cpu:
        from oracle.twin import Twin
        # every core this process may run on (the GPU box allots 16 of its host cores to a
        # one-GPU job); os.cpu_count() is reported beside it
        ncores = len(os.sched_getaffinity(0))
        grid = system.grid.to_reference_grid()
        ffa = [t.cpu().numpy() for t in (system.ff.rc, system.ff.ac, system.ff.ivor, system.ff.emin,
                                         system.ff.rmin2)]
        twins = [Twin(grid, ffa, w.coor_rec, w.rec_types - 1, l.coor, l.types - 1, l.types - 1) for l in w.ligands]
        # calibrate: 1 pose of each ligand, then scale the sample to ~cpu_seconds
        t0 = time.perf_counter()
        nst_cal = 10
        for k, tw in enumerate(twins):
            sl = int(plan.offsets[k])
            tw.minimize(np.concatenate([w.quat[sl:sl + 1], w.trans[sl:sl + 1]], 1), nst_cal, w.lr_rot, w.lr_trans,
                        nthreads=1)
        per_ps = (time.perf_counter() - t0) / (len(twins) * nst_cal)
        per_lig = max(1, int(args.cpu_seconds * ncores / (per_ps * w.nsteps * len(twins))))
        per_lig = min(per_lig, w.pose_counts[0])
        t0 = time.perf_counter()
        cps = 0
        for k, tw in enumerate(twins):
            sl = slice(int(plan.offsets[k]), int(plan.offsets[k]) + per_lig)
            tw.minimize(np.concatenate([w.quat[sl], w.trans[sl]], 1), w.nsteps, w.lr_rot, w.lr_trans,
                        nthreads=ncores)
            cps += per_lig * w.nsteps
        tc = time.perf_counter() - t0
        cpu_model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        result["cpu_baseline"] = {
            "value": cps / tc, "unit": "pose-steps/s", "cores": ncores, "kind": "port",
            "host_cpu_count": os.cpu_count(), "cpu_model": cpu_model,
            "sample": f"FP32 C/OpenMP twin of the same algorithm on all {ncores} cores this process may use "
                      f"(os.cpu_count() {os.cpu_count()}): {per_lig} poses of each of the {len(twins)} ligands x "
                      f"{w.nsteps} ADAM steps ({tc:.1f}s), {cpu_model}"}
        log(f"[bench] cpu twin: {cps / tc:.3e} pose-steps/s on {ncores} threads over {tc:.1f}s")

    if not args.no_scoring:
        result["scoring"] = scoring_rates(dev, system.launch.waves_per_block)

    if args.serve_jobs:
        from attract_jax_amd import workloads as W
        from attract_jax_amd.serve import GpuBackend, Scheduler, run_stream
        t0 = time.time()
        specs, receptors = W.c5_stream(args.serve_jobs, rate=args.serve_rate)
        log(f"[bench] serving workload generated in {time.time() - t0:.1f}s ({len(receptors)} receptors)")
        be = GpuBackend(dev, receptors, waves_per_block=system.launch.waves_per_block,
                        launch_options=system.launch.options & (_lib.ATJ_LAUNCH_CLAIM_MASK | _lib.ATJ_LAUNCH_TAIL_MASK |
                                                                    _lib.ATJ_LAUNCH_TAIL_ROUNDS_MASK))
        budget = int(args.serve_budget_gib * 2 ** 30)
        sch = Scheduler(be, hbm_budget=budget, chunk_steps=25, chunk_work=1_000_000)
        # host time of each scheduler tick (admission builds grids on the host thread) and the
        # caching allocator's retries (a retry frees every cached block, synchronizing the device)
        ticks, tick = [], sch.tick

        def timed_tick(now=None):
            t = time.perf_counter()
            tick(now)
            ticks.append(time.perf_counter() - t)

        sch.tick = timed_tick
        host_ops = {}

        def timed_op(name, fn):
            def run(*a, **k):
                t = time.perf_counter()
                r = fn(*a, **k)
                d = time.perf_counter() - t
                n, tot, mx = host_ops.get(name, (0, 0.0, 0.0))
                host_ops[name] = (n + 1, tot + d, max(mx, d))
                return r
            return run

        for name in ("make_group", "load_job", "launch", "fetch_job", "drop_group", "grow_group", "reset_group"):
            setattr(be, name, timed_op(name, getattr(be, name)))
        mem0 = torch.cuda.memory_stats(dev)
        torch.cuda.reset_peak_memory_stats(dev)
        lat, makespan, ps = run_stream(sch, specs, timeout=600)
        mem1 = torch.cuda.memory_stats(dev)
        waits = sch.admission_waits()
        nposes = [len(sp.quat) for sp in specs]
        plain_gb = [r.plain_bytes() / 1e9 for r in receptors.values()]
        corner = sum(1 for lay in be.layouts if lay == _lib.ATJ_LAYOUT_CORNER)
        result["serving"] = {
            "workload": "C5: Poisson arrivals, every job with its own receptor (500-4000 beads, own pair table) "
                        "whose grid is built on the device at admission, ligand 5-60 beads, 1e3-2e5 poses, "
                        "100-300 ADAM steps; latency = completion - arrival",
            "jobs": len(specs), "completed": int(sum(j.state == "done" for j in sch.done)),
            "arrival_rate_jobs_per_s": args.serve_rate, "hbm_budget_gib": args.serve_budget_gib,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "p50_admission_wait_s": float(np.percentile(waits, 50)),
            "p99_admission_wait_s": float(np.percentile(waits, 99)), "evictions": sch.evictions,
            "groups_corner_layout": corner, "groups_plain_layout": len(be.layouts) - corner,
            "grid_build_s_p50": float(np.percentile(be.build_seconds, 50)),
            "grid_build_s_max": float(np.max(be.build_seconds)),
            "tick_s_p99": float(np.percentile(ticks, 99)), "tick_s_max": float(np.max(ticks)),
            "allocator_retries": int(mem1.get("num_alloc_retries", 0) - mem0.get("num_alloc_retries", 0)),
            "reserved_gib_start": mem0.get("reserved_bytes.all.current", 0) / 2 ** 30,
            "reserved_gib_peak": mem1.get("reserved_bytes.all.peak", 0) / 2 ** 30,
            "host_ops": {k: {"calls": n, "total_s": round(t, 4), "max_s": round(m, 4)}
                         for k, (n, t, m) in host_ops.items()},
            "plain_grid_gb": [round(min(plain_gb), 3), round(max(plain_gb), 3)],
            "makespan_s": makespan, "pose_steps": ps, "aggregate_pose_steps_per_s": ps / makespan,
            "poses_per_job": [int(min(nposes)), int(max(nposes))], "chunk_steps": sch.chunk_steps,
            "chunk_work": sch.chunk_work}
        log(f"[bench] serving: p50 {result['serving']['p50_latency_s']:.3f}s p99 "
            f"{result['serving']['p99_latency_s']:.3f}s, admission wait p99 "
            f"{result['serving']['p99_admission_wait_s']:.3f}s, {sch.evictions} evictions, {corner} corner / "
            f"{len(be.layouts) - corner} plain grids, {ps / makespan:.3e} pose-steps/s over {makespan:.1f}s; "
            f"longest tick {np.max(ticks):.3f}s, allocator retries {result['serving']['allocator_retries']}, "
            f"reserved {result['serving']['reserved_gib_start']:.1f} -> peak "
            f"{result['serving']['reserved_gib_peak']:.1f} GiB")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
