"""Differentiable pose energies: torch.autograd.Function over the HIP kernels.

Forward: energies (N,) of poses M (N, 4, 4) (atj_energrad).  Backward:
  dL/dM         = dL/dE[:, None, None] * dE/dM        (kept from the forward launch)
  dL/d(rc, ac, emin) = sum_pose dL/dE_pose * dE_pose/d(rc, ac, emin)
                  (atj_energrad_adjoint: recomputes the pairs, nothing per pair is stored)
The pair-table gradient covers the neighbour-list term (the potential grids are data);
rmin2 and ivor get no gradient (E is continuous in rmin2 when emin, rmin2 are the
reference's derived values, test-grid-energrad.py:71-72, and ivor is discrete).
Parameters (rmin, emin) map to (rc, ac, emin, rmin2) with differentiable torch ops
(pair_table_from_rmin_emin), so autograd chains down to them.

Gradient convention: flags with ATJ_GRAD_EXACT give the exact derivative of the
energy (finite-difference checkable); without it the ATTRACT convention of the
reference (test-grid-energrad.py:129-147, 232-239).
"""
import ctypes
import math

import torch

from . import _lib
from .device import ForceField, _ptr, _stream


def pair_table_from_rmin_emin(rmin, emin_abs, ivor):
    """Differentiable ATTRACT parameters from (rmin, |emin|): rbc = rmin sqrt(3)/2,
    abc = 256/27 |emin|, rc = abc rbc^8, ac = abc rbc^6; emin = -27 ac^4/(256 rc^3),
    rmin2 = 4 rc/(3 ac) (attract-original/convert-attract-par.py:16-22, test-grid-energrad.py:71-72)."""
    rbc = rmin * (math.sqrt(3.0) / 2.0)
    abc = (256.0 / 27.0) * emin_abs.abs()
    rc = abc * rbc ** 8
    ac = abc * rbc ** 6
    emin = -27.0 * ac ** 4 / (256.0 * rc ** 3)
    rmin2 = 4.0 * rc / (3.0 * ac)
    return rc, ac, ivor, emin, rmin2


class PoseEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mat, rc, ac, ivor, emin, rmin2, system, ligand, flags, blocks):
        lib = _lib.load()
        ff = ForceField.from_tensors(rc, ac, ivor, emin, rmin2)
        M = mat.detach().to(torch.float32).contiguous().reshape(-1, 16)
        N = M.shape[0]
        d = system.device
        energy = torch.empty(N, dtype=torch.float32, device=d)
        dEdM = torch.empty((N, 16), dtype=torch.float32, device=d)
        jobs = system.jobs_tensor([ligand])
        _lib.check(lib.atj_energrad(ctypes.byref(system._gs), ctypes.byref(ff.struct()), ctypes.byref(system._rs),
                                    _ptr(jobs), 1, ctypes.c_void_p(0), ctypes.c_void_p(0), _ptr(M), N,
                                    _ptr(energy), _ptr(dEdM), flags, ctypes.byref(system.launch),
                                    _ptr(system.counter), ctypes.c_void_p(0), _stream(d)))
        ctx.save_for_backward(M, dEdM, ff.rc, ff.ac, ff.ivor, ff.emin, ff.rmin2)
        ctx.system, ctx.ligand, ctx.flags, ctx.blocks = system, ligand, flags, blocks
        ctx.jobs = jobs
        ctx.mat_dtype = mat.dtype
        return energy

    @staticmethod
    def backward(ctx, gE):
        M, dEdM, rc, ac, ivor, emin, rmin2 = ctx.saved_tensors
        gE = gE.detach().to(torch.float32).contiguous()
        grad_mat = None
        if ctx.needs_input_grad[0]:
            grad_mat = (gE[:, None] * dEdM).reshape(-1, 4, 4).to(ctx.mat_dtype)
        g_rc = g_ac = g_emin = None
        if any(ctx.needs_input_grad[1:6]):
            g = param_adjoint(ctx.system, ctx.ligand, M, ForceField(rc, ac, ivor, emin, rmin2), gE, ctx.flags,
                              ctx.blocks, ctx.jobs)
            g_rc, g_ac, g_emin = g[0], g[1], g[2]
        return grad_mat, g_rc, g_ac, None, g_emin, None, None, None, None, None


def param_adjoint(system, ligand, M, ff, weights, flags, blocks=None, jobs=None, waves_per_block=None,
                  options=None):
    """sum_pose w_pose dE_pose/d(rc, ac, emin) -> (3, nrt, nlt) float32 (neighbour-list term).
    Reproducible bit for bit for a given launch shape (blocks, waves per block)."""
    lib = _lib.load()
    d = system.device
    N = M.shape[0]
    if blocks is None:  # one block per CU: a block's gradient tables fill most of its LDS
        blocks = torch.cuda.get_device_properties(d).multi_processor_count
    ntp = ff.nrt * ff.nlt
    partial = torch.empty((blocks, 3 * ntp), dtype=torch.float32, device=d)
    out = torch.empty(3 * ntp, dtype=torch.float32, device=d)
    if jobs is None:
        jobs = system.jobs_tensor([ligand])
    launch = _lib.AtjLaunch(waves_per_block or 16, blocks,
                            system.launch.options if options is None else options)
    _lib.check(lib.atj_energrad_adjoint(ctypes.byref(system._gs), ctypes.byref(ff.struct()),
                                        ctypes.byref(system._rs), _ptr(jobs), 1, ctypes.c_void_p(0), _ptr(M), N,
                                        _ptr(weights), ctypes.c_void_p(0), ctypes.c_void_p(0), _ptr(partial), flags,
                                        ctypes.byref(launch), ctypes.c_void_p(0), _stream(d)))
    _lib.check(lib.atj_reduce_partials(_ptr(partial), blocks, 3 * ntp, _ptr(out), _stream(d)))
    return out.reshape(3, ff.nrt, ff.nlt)


def pose_energy(mat, rc, ac, ivor, emin, rmin2, system, ligand, flags=None, blocks=None):
    """Differentiable energies (N,) of poses mat (N,4,4) w.r.t. mat and (rc, ac, emin)."""
    if flags is None:
        flags = system.flags
    return PoseEnergy.apply(mat, rc, ac, ivor, emin, rmin2, system, ligand, flags, blocks)


def dof_gradient(mat, dEdM):
    """(dE/dt, dE/domega) (N, 6) from dE/dM, omega = world-frame rotation vector about the
    ligand centre: dE/domega_a = eps_abc (R^T A)_bc with A = dE/dM[0:3, 0:3]."""
    R = mat[:, :3, :3]
    A = dEdM[:, :3, :3]
    B = (R.unsqueeze(3) * A.unsqueeze(2)).sum(dim=1)
    tau = torch.stack([B[:, 1, 2] - B[:, 2, 1], B[:, 2, 0] - B[:, 0, 2], B[:, 0, 1] - B[:, 1, 0]], dim=1)
    return torch.cat([dEdM[:, 3, :3], tau], dim=1)
