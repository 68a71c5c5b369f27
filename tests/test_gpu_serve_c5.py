"""Serving the C5 way on the GPU, bounded: every job brings its own receptor (500-1200 beads, own
pair table), whose grid is built on the device when the job is admitted; an HBM budget of about
one and a half corner-packed grids makes jobs wait for memory, falls back to the plain layout
and evicts idle grids.  Every job's result is the same, bit for bit, as when it runs alone with
the device to itself, and a job's poses match the oracle's FP32 twin."""
import numpy as np
import pytest
import torch

from attract_jax_amd import _lib
from attract_jax_amd import workloads as W
from attract_jax_amd.device import DeviceGrid, ForceField, Receptor
from attract_jax_amd.serve import GpuBackend, Scheduler, run_stream
from oracle.twin import Twin
from tests.systems import pose_rmsd, twin_conditioning

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def stream():
    specs, receptors = W.c5_stream(njobs=8, rate=40.0, seed=4242, pose_scale=0.02, nrec_range=(500, 1201))
    for sp in specs:
        sp.nsteps = min(sp.nsteps, 120)
    return specs, receptors


def run(specs, receptors, budget, corner_min_work=1e5):
    be = GpuBackend(DEV, receptors, waves_per_block=8, corner_min_work=corner_min_work)
    s = Scheduler(be, hbm_budget=budget, chunk_steps=20, initial_slots=1024)
    lat, makespan, ps = run_stream(s, specs, timeout=240, stall=60)
    return s, be, lat


def test_c5_stream_under_memory_pressure_equals_jobs_alone():
    specs, receptors = stream()
    corner = max(r.corner_bytes() + r.list_bytes() for r in receptors.values())
    s, be, lat = run(specs, receptors, budget=int(1.5 * corner))
    assert all(j.state == "done" for j in s.done) and len(s.done) == len(specs)
    print(f"C5 bounded: {len(specs)} jobs, evictions {s.evictions}, layouts {be.layouts}, admission waits "
          f"{np.round(s.admission_waits(), 3).tolist()}, latencies {np.round(lat, 3).tolist()}")
    assert s.evictions >= 1
    assert len(set(be.layouts)) == 2  # both layouts were used
    mixed = {j.id: j.result for j in s.jobs}
    for i, sp in enumerate(stream()[0]):
        sp.arrival = 0.0
        alone, _, _ = run([sp], {sp.receptor_key: receptors[sp.receptor_key]}, budget=64 * 2 ** 30)
        res = alone.jobs[0].result
        assert np.array_equal(res["pose"], mixed[i]["pose"]), i
        assert np.array_equal(res["energy"], mixed[i]["energy"]), i
        assert np.isfinite(res["energy"]).all()


def test_c5_job_matches_twin():
    specs, receptors = stream()
    sp = max(specs, key=lambda x: len(x.quat))
    sp.arrival = 0.0
    r = receptors[sp.receptor_key]
    s, be, _ = run([sp], {r.key: r}, budget=64 * 2 ** 30)
    res = s.jobs[0].result
    ff = ForceField.create(r.rc, r.ac, r.ivor, device=DEV)
    rec = Receptor.create(r.coor, r.types - 1, DEV)
    grid = DeviceGrid.build(rec, ff, np.arange(W.NTYPES), layout=_lib.ATJ_LAYOUT_PLAIN, device=DEV, keep_plain=True,
                            **r.grid_kw).to_reference_grid()
    ffa = [t.cpu().numpy() for t in (ff.rc, ff.ac, ff.ivor, ff.emin, ff.rmin2)]
    tw = Twin(grid, ffa, r.coor, r.types - 1, sp.ligand_coor, sp.ligand_types - 1, sp.ligand_types - 1)
    Pt, Et, good, _ = twin_conditioning(tw, np.concatenate([sp.quat, sp.trans], 1), sp.nsteps, sp.lr_rot,
                                        sp.lr_trans, sp.ligand_coor)
    rr = pose_rmsd(res["pose"], Pt, sp.ligand_coor)
    de = np.abs(res["energy"] - Et) / np.maximum(np.abs(Et), 1.0)
    print(f"C5 served job ({len(sp.quat)} poses, {sp.nsteps} steps, {r.coor.shape[0]}-bead receptor) vs the twin: "
          f"{good.sum()} of {len(good)} poses well conditioned; RMSD max {rr[good].max():.1e} median "
          f"{np.median(rr[good]):.1e} A; energy rel max {de[good].max():.1e} p99 {np.percentile(de[good], 99):.1e} "
          f"median {np.median(de[good]):.1e}")
    assert good.mean() >= 0.9
    assert rr[good].max() < 0.05 and np.median(rr) < 0.01
    assert np.percentile(de[good], 99) < 1e-4
