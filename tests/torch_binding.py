"""Thin torch.autograd binding of the GPU projection, for the adjoint tests only.

forward:  (u_proj, v_proj) = P(u, v; beta_x, beta_y), the pressure kept in the
          projection context as the only checkpoint;
backward: one adjoint solve through the C ABI (p2d_adjoint) gives the
          gradients with respect to u, v, beta_x and beta_y (zero on the wall
          faces: wall velocities and wall coefficients are boundary data).
"""
import torch


class Projection(torch.autograd.Function):
    @staticmethod
    def forward(fctx, p2d_ctx, phi, jc, u, v, bx, by, dt):
        p2d_ctx.set_coefficients(bx.detach().contiguous(), by.detach().contiguous(), phi, jc)
        up = u.detach().clone().contiguous()
        vp = v.detach().clone().contiguous()
        p2d_ctx.project(up, vp, dt)
        fctx.p2d = p2d_ctx
        return up, vp

    @staticmethod
    def backward(fctx, gu, gv):
        du, dv, dbx, dby, _ = fctx.p2d.adjoint(gu.contiguous(), gv.contiguous())
        return None, None, None, du, dv, dbx, dby, None


def project(p2d_ctx, phi, jc, u, v, bx, by, dt=1.0):
    return Projection.apply(p2d_ctx, phi, jc, u, v, bx, by, dt)
