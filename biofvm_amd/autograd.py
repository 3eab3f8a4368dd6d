"""torch.autograd bridge to the adjoint (the bfvm_tape_* entry points of the
C ABI), used for gradient checks and the fitting driver (``biofvm_amd.fit``).

    model = Model(case, precision=FP32)          # grid, fixed inputs, one handle
    rho_N = LodSteps.apply(rho0, D, lam, secretion, uptake, model)
    loss = 0.5 * ((rho_N - target) ** 2).sum()
    loss.backward()                              # runs bfvm_tape_backward

Tensors are float64 on the host; the handle moves them through the C ABI,
and every step, forward and adjoint, runs in the library's HIP kernels.  The
fixed inputs of ``case`` (grid, dt, steps, phases, agent positions, volumes
and saturation densities, Dirichlet nodes, bulk fields) are set once.
"""
import numpy as np
import torch

from .device import DeviceMicroenvironment
from ._capi import FP64


class Model:
    """One handle plus the fixed inputs of a case (see ``workloads.make_case``)."""

    def __init__(self, case, precision=FP64, checkpoint_every=50):
        self.case = case
        self.checkpoint_every = int(checkpoint_every)
        self.env = DeviceMicroenvironment(case["nx"], case["ny"], case["nz"], case["S"], case["dx"],
                                          precision)
        if case.get("lod_dims", 3) != 3:
            self.env.set_lod_dimensions(case["lod_dims"])
        if case.get("dir_idx") is not None and len(case["dir_idx"]) > 0:
            self.env.set_dirichlet(case["dir_idx"], case["dir_val"], case.get("dir_mask"))
        if case.get("bulk_supply") is not None:
            self.env.set_bulk(case["bulk_supply"], case["bulk_target"], case["bulk_uptake"])
        self.n_agents = 0 if case.get("agent_volume") is None else len(case["agent_volume"])
        self.generation = 0

    def close(self):
        self.env.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self):
        """The case's own parameters as float64 tensors (rho0, D, lambda, secretion, uptake)."""
        c, S = self.case, self.case["S"]
        sec = c["agent_secretion"] if self.n_agents else np.zeros(0)
        upt = c["agent_uptake"] if self.n_agents else np.zeros(0)
        return (torch.tensor(c["rho0"]), torch.tensor(c["D"]), torch.tensor(c["lam"]),
                torch.tensor(sec).reshape(-1, S), torch.tensor(upt).reshape(-1, S))


class LodSteps(torch.autograd.Function):
    """rho_N = the case's `steps` steps from rho0 with D, lambda and the
    agents' secretion / uptake rates; backward through the adjoint."""

    @staticmethod
    def _record(model, rho0, D, lam, secretion, uptake):
        c, m = model.case, model.env
        m.set_coefficients(D.detach().numpy(), lam.detach().numpy())
        if model.n_agents:
            m.set_agent_rates(c["agent_pos"], c["agent_volume"], secretion.detach().numpy().reshape(-1),
                              c["agent_saturation"], uptake.detach().numpy().reshape(-1), c["dt"])
        m.upload(rho0.detach().numpy())
        m.tape_record(c["dt"], c["steps"], c["phases"], model.checkpoint_every)
        model.generation += 1
        return model.generation

    @staticmethod
    def forward(ctx, rho0, D, lam, secretion, uptake, model):
        ctx.generation = LodSteps._record(model, rho0, D, lam, secretion, uptake)
        ctx.model = model
        ctx.shapes = (secretion.shape, uptake.shape)
        ctx.save_for_backward(rho0, D, lam, secretion, uptake)
        return torch.from_numpy(model.env.download())

    @staticmethod
    def backward(ctx, grad_out):
        model = ctx.model
        if model.generation != ctx.generation:
            # the model ran forward since (another output): record this run again
            ctx.generation = LodSteps._record(model, *ctx.saved_tensors)
        m = model.env
        m.tape_seed(grad_out.detach().double().numpy())
        g = m.tape_backward(model.n_agents)
        gs = torch.from_numpy(g["grad_secretion"]) if model.n_agents else torch.zeros(ctx.shapes[0],
                                                                                     dtype=torch.float64)
        gu = torch.from_numpy(g["grad_uptake"]) if model.n_agents else torch.zeros(ctx.shapes[1],
                                                                                  dtype=torch.float64)
        return (torch.from_numpy(g["grad_rho0"]), torch.from_numpy(g["grad_D"]),
                torch.from_numpy(g["grad_lambda"]), gs.reshape(ctx.shapes[0]), gu.reshape(ctx.shapes[1]),
                None)
