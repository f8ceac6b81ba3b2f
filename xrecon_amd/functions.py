"""Autograd adjoint pairs (the reference's ``xrecon/projector/functions.py:96-302``).

Every operator is linear, and its backward is its exact transpose computed by the
hand-written adjoint kernel of the same plan: A <-> A^T, B <-> B^T, B_W <-> B_W^T.  The
backward is itself built from the same autograd functions, so higher-order gradients
(e.g. Hessian-vector products through an unrolled reconstruction) work too.

``LinearOp`` is the plan-level function used by the projector modules; the
reference-named classes (``Projection3d.apply(input, options, angles)`` etc.) resolve a
plan from the options/angles tensors through the plan cache and then behave the same.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from . import _backend as be


class LinearOp(Function):
    """y = op(x) for a plan and an operator id; dL/dx = adjoint(op)(dL/dy)."""

    @staticmethod
    def forward(ctx, x, plan, op):
        x = x.contiguous()
        ctx.plan = plan
        ctx.op = op
        return plan.execute(op, x, be.current_stream(x))

    @staticmethod
    def backward(ctx, grad):
        return LinearOp.apply(grad.contiguous(), ctx.plan, be.ADJOINT[ctx.op]), None, None


def apply_op(plan, op: int, x: torch.Tensor) -> torch.Tensor:
    return LinearOp.apply(x, plan, op)


def _make(name: str, op: int, is2d: bool, doc: str):
    def forward(ctx, input, option, angles):
        plan = be.get_plan(option, angles, input, is2d)
        ctx.save_for_backward(option, angles)
        ctx.plan = plan
        return plan.execute(op, input.contiguous(), be.current_stream(input))

    def backward(ctx, grad_output):
        return LinearOp.apply(grad_output.contiguous(), ctx.plan, be.ADJOINT[op]), None, None

    cls = type(name, (Function,), {"forward": staticmethod(forward), "backward": staticmethod(backward),
                                   "__doc__": doc, "__module__": __name__})
    return cls


Projection2d = _make("Projection2d", be.OP_PROJECTION, True, "2D A; backward A^T (functions.py:96-111).")
BackProjection2d = _make("BackProjection2d", be.OP_BACKPROJECTION, True, "2D B; backward B^T (:114-129).")
WeightedBackProjection2d = _make("WeightedBackProjection2d", be.OP_BACKPROJECTION_W, True,
                                 "2D FBP-weighted B; backward B_W^T (:131-146).")
ProjectionTranspose2d = _make("ProjectionTranspose2d", be.OP_PROJECTION_T, True, "2D A^T; backward A (:149-164).")
BackProjectionTranspose2d = _make("BackProjectionTranspose2d", be.OP_BACKPROJECTION_T, True,
                                  "2D B^T; backward B (:167-182).")
WeightedBackProjectionTranspose2d = _make("WeightedBackProjectionTranspose2d", be.OP_BACKPROJECTION_W_T, True,
                                          "2D B_W^T; backward B_W (:185-200).")

Projection3d = _make("Projection3d", be.OP_PROJECTION, False, "3D A; backward A^T (functions.py:203-217).")
ProjectionTranspose3d = _make("ProjectionTranspose3d", be.OP_PROJECTION_T, False, "3D A^T; backward A (:220-234).")
BackProjection3d = _make("BackProjection3d", be.OP_BACKPROJECTION, False, "3D B; backward B^T (:237-251).")
BackProjectionTranspose3d = _make("BackProjectionTranspose3d", be.OP_BACKPROJECTION_T, False,
                                  "3D B^T; backward B (:254-268).")
WeightedBackProjection3d = _make("WeightedBackProjection3d", be.OP_BACKPROJECTION_W, False,
                                 "3D FDK-weighted B; backward B_W^T (:271-285).")
WeightedBackProjectionTranspose3d = _make("WeightedBackProjectionTranspose3d", be.OP_BACKPROJECTION_W_T, False,
                                          "3D B_W^T; backward B_W (:288-302).")
