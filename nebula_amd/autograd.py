"""Differentiable rollouts of the MPM engine for torch.

``rollout(engine, state, params, n_substeps, dt)`` runs ``n_substeps``
substeps of ``engine`` from ``state`` as a recorded run
(nb_mpm_forward_recorded) and returns the final state; its backward calls
nb_mpm_backward, which walks the run in reverse with checkpoints every
``segment`` substeps.  The reference has no differentiation: this binding
sits on this library's own reverse-mode entry points (include/nebula_mpm.h).

Layouts:
  state   (8, n, 4) float32 CUDA tensor: the engine's eight float4 arrays
          P (x, Vp), v (v, material id bits), F0..2 (F columns, tau diagonal
          in w), B0..2 (B columns, tau off-diagonals in w) back to back.
          The material id lane (state[1, :, 3]) gets no gradient.
  params  (material_count, 4) tensor: (mu, lambda, M_friction, q_cohesion)
          of each material; the other material fields come from the table
          last given to ``engine.setMaterials``.  ``material_params_from``
          builds it from Young's modulus, Poisson's ratio, friction angle and
          cohesion, so gradients reach those through torch's chain rule.

The engine must be in the adjoint's domain (corrected_svd = True; PIC or
APIC; elastic, Von Mises or Drucker-Prager materials; no analytic
colliders).  One recorded run is held per engine: the backward of a rollout
must come before the engine records another one.  The engine keeps the
recording's memory until the backward (or ``engine.adjointEnd()``).

Import this module (or torch) before the engine library loads: the library
then binds to torch's HIP runtime, and tensors' device pointers are valid for
it.  Importing it after the library loaded without torch raises, because the
process then holds two HIP runtimes (``mpm.check_single_hip_runtime``).
"""
import itertools

import torch

from . import mpm

if mpm._lib is not None:
    mpm.check_single_hip_runtime()

_record_ids = itertools.count(1)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, params, engine, n_substeps, dt, segment, tapes):
        if state.dim() != 3 or state.shape[0] != 8 or state.shape[2] != 4 or state.dtype != torch.float32:
            raise ValueError("state must be a (8, n, 4) float32 tensor")
        if not state.is_cuda:
            raise ValueError("state must live on the engine's GPU")
        mats = engine.materials()
        if params.shape != (len(mats), 4):
            raise ValueError(f"params must be ({len(mats)}, 4): (mu, lambda, M_friction, q_cohesion) per material")
        for m, row in zip(mats, params.detach().double().cpu().tolist()):
            m.mu, m.lambda_, m.M_friction, m.q_cohesion = row
        engine.setMaterials(mats)
        st = state.detach().contiguous()
        n = st.shape[1]
        torch.cuda.current_stream(st.device).synchronize()
        engine.setStateFrom(st.data_ptr(), n)
        engine.adjointBegin(n_substeps, segment, tapes)
        engine.forwardRecorded(dt)
        out = torch.empty_like(st)
        engine.getStateTo(out.data_ptr())
        rid = next(_record_ids)
        engine._record_id = rid
        ctx.engine, ctx.record_id, ctx.n = engine, rid, n
        ctx.params_meta = (params.dtype, params.device)
        return out

    @staticmethod
    def backward(ctx, g_out):
        e = ctx.engine
        if getattr(e, "_record_id", None) != ctx.record_id:
            raise RuntimeError("the engine recorded another run after this rollout; its backward is gone")
        g = g_out.detach().to(torch.float32).contiguous().clone()
        g[1, :, 3] = 0.0  # material id bits
        n = ctx.n
        torch.cuda.current_stream(g.device).synchronize()
        for i in range(8):
            e.setAdjointField(i, g.data_ptr() + i * n * 16, n)
        dpar = e.backward()[: len(e.materials())]
        g_state = torch.empty_like(g)
        for i in range(8):
            e.getAdjointField(i, g_state.data_ptr() + i * n * 16, n)
        e.adjointEnd()
        e._record_id = None
        dtype, device = ctx.params_meta
        g_params = torch.as_tensor(dpar, dtype=dtype, device=device)
        return g_state, g_params, None, None, None, None, None


def rollout(engine: "mpm.MPMEngine", state: torch.Tensor, params: torch.Tensor, n_substeps: int, dt: float,
            segment: int = 0, tapes: int = 1) -> torch.Tensor:
    """Final state after ``n_substeps`` substeps of size ``dt`` from ``state``
    with material parameters ``params``; differentiable in both.  segment = 0
    keeps every substep's state (no recompute in the backward); a smaller
    segment trades memory for one recomputed forward per segment, except for
    the ``tapes`` segments whose states are kept at once (0: as many as fit in
    free device memory).  The gradients do not depend on segment or tapes."""
    seg = n_substeps if segment <= 0 else segment
    return _Rollout.apply(state, params, engine, int(n_substeps), float(dt), int(seg), int(tapes))


def material_params(engine: "mpm.MPMEngine", device=None) -> torch.Tensor:
    """(material_count, 4) float32 tensor of the engine's (mu, lambda,
    M_friction, q_cohesion), a starting point for ``params``."""
    rows = [[m.mu, m.lambda_, m.M_friction, m.q_cohesion] for m in engine.materials()]
    return torch.tensor(rows, dtype=torch.float32, device=device)


def material_params_from(E: torch.Tensor, nu: torch.Tensor, friction_angle: torch.Tensor,
                         cohesion: torch.Tensor) -> torch.Tensor:
    """(n, 4) rows (mu, lambda, M_friction, q_cohesion) from per-material
    Young's modulus E, Poisson's ratio nu, friction angle (radians) and
    cohesion, as the reference's presets derive them: mu = E / (2 (1 + nu)),
    lambda = E nu / ((1 + nu)(1 - 2 nu)) (src/core/MaterialParams.h:42-43) and
    M = tan(friction angle) (:22).  Differentiable, so a rollout's gradient
    with respect to params becomes one with respect to (E, nu, angle,
    cohesion)."""
    E, nu, phi, qc = torch.broadcast_tensors(E, nu, friction_angle, cohesion)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return torch.stack([mu, lam, torch.tan(phi), qc], dim=-1)
