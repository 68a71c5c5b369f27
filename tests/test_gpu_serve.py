"""Serving on the GPU: overlapping jobs on two receptors run through the continuous-batching
loop; every job's result equals the result of the same job run alone (bit for bit: a
pose's arithmetic does not depend on the batch it shares a launch with), and matches
the oracle's CPU twin within the minimization tolerance."""
import numpy as np
import pytest
import torch

from attract_jax_amd import workloads as W
from attract_jax_amd.serve import GpuBackend, JobSpec, Scheduler, run_stream

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def small_receptors():
    rs = []
    for i, n in enumerate((300, 500)):
        rng = np.random.default_rng(70 + i)
        radius = 25.0 * (n / 2000.0) ** (1 / 3)
        coor = W.random_receptor(n, radius, 2.25, rng)
        rc, ac, ivor, _, _ = W.random_pair_table(rng)
        d = int(np.ceil(2 * (radius + 12.0) / 0.9)) + 1
        o = -0.9 * (d - 1) / 2
        rs.append(dict(key=i, nrec=n, radius=radius, coor=coor, types=rng.integers(1, 33, n), rc=rc, ac=ac,
                       ivor=ivor, grid_kw=dict(origin=(o, o, o), spacing=0.9, dim=(d, d, d), dim2=(32, 32, 32),
                                               gridextension=16, plateaudis=6.0, neighbourdis=6.8)))
    return rs


def job_specs(rs):
    specs = []
    for i, (key, nb, n, steps, t) in enumerate([(0, 7, 900, 60, 0.0), (1, 21, 400, 40, 0.0),
                                                 (0, 33, 300, 45, 0.01), (1, 5, 1200, 30, 0.02),
                                                 (0, 64, 150, 35, 0.02)]):
        rng = np.random.default_rng(100 + i)
        r = rs[key]
        q, tr = W.posemod.random_poses(n, rng, r["radius"], r["radius"] + 6)
        specs.append(JobSpec(key, W.random_chain(nb, rng), rng.integers(1, 33, nb), q, tr, steps, arrival=t))
    return specs


def run(rs, specs, chunk):
    be = GpuBackend(DEV, W.c5_backend_receptors(rs, DEV), waves_per_block=8)
    s = Scheduler(be, hbm_budget=40 * 2 ** 30, chunk_steps=chunk, initial_slots=2048)
    lat, makespan, ps = run_stream(s, specs, timeout=120)
    return s, lat


def test_same_result_alone_and_in_the_mix():
    rs = small_receptors()
    specs = job_specs(rs)
    s_mix, lat = run(rs, specs, chunk=10)
    assert len(s_mix.done) == len(specs) and all(j.state == "done" for j in s_mix.done)
    assert np.all(lat > 0)
    mixed = {j.id: j.result for j in s_mix.jobs}
    for i, sp in enumerate(job_specs(rs)):
        sp.arrival = 0.0
        s_one, _ = run(rs, [sp], chunk=25)
        alone = s_one.jobs[0].result
        assert np.array_equal(alone["pose"], mixed[i]["pose"]), i
        assert np.array_equal(alone["energy"], mixed[i]["energy"]), i
        assert np.isfinite(alone["energy"]).all()


def test_served_job_matches_twin():
    from oracle.twin import Twin
    rs = small_receptors()
    sp = job_specs(rs)[1]
    sp.arrival = 0.0
    s, _ = run(rs, [sp], chunk=15)
    res = s.jobs[0].result
    r = rs[sp.receptor_key]
    system = W.c5_backend_receptors(rs, DEV, keep_plain=True)[sp.receptor_key]["build"]()
    grid = system.grid.to_reference_grid()
    ffa = [t.cpu().numpy() for t in (system.ff.rc, system.ff.ac, system.ff.ivor, system.ff.emin, system.ff.rmin2)]
    tw = Twin(grid, ffa, r["coor"], r["types"] - 1, sp.ligand_coor, sp.ligand_types - 1, sp.ligand_types - 1)
    P, E = tw.minimize(np.concatenate([sp.quat, sp.trans], 1), sp.nsteps, sp.lr_rot, sp.lr_trans)
    d = np.abs(res["pose"][:, 4:] - P[:, 4:]).max(axis=1)
    assert np.median(d) < 1e-3 and (d < 0.05).mean() >= 0.95
