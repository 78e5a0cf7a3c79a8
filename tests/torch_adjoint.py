"""torch.autograd.Function around the adjoint entry points.  TEST HARNESS
ONLY: it lets a loss written in torch drive pn1_adjoint_backward, so the
tests can compare the library's gradients with autograd's own finite
difference checks.  The computation is libpn1's; torch holds CPU float64
tensors and does the loss."""
import numpy as np
import torch


class Trajectory(torch.autograd.Function):
    """(x0, v0, m0) -> (x_T, v_T) after `steps` steps of the stage driver on
    a pn_amd.lib.Context whose eps, c, dt are set; `m_rest` holds the star
    masses (m_rest[0] is replaced by m0)."""

    @staticmethod
    def forward(fctx, x0, v0, m0, pnctx, m_rest, steps, interval):
        m = np.array(m_rest, dtype=np.float64)
        m[0] = float(m0)
        pnctx.upload(x0.detach().numpy(), v0.detach().numpy(), m)
        pnctx.adjoint_forward(int(steps), int(interval))
        xt, vt = pnctx.download()
        fctx.pnctx = pnctx
        return torch.from_numpy(xt), torch.from_numpy(vt)

    @staticmethod
    def backward(fctx, gxt, gvt):
        gx0, gv0, gm0 = fctx.pnctx.adjoint_backward(gxt.contiguous().numpy(),
                                                    gvt.contiguous().numpy())
        return (torch.from_numpy(gx0), torch.from_numpy(gv0),
                torch.tensor(gm0, dtype=torch.float64), None, None, None, None)


def radial_loss_torch(x, v, r_cut):
    """Mean squared radial velocity (relative to the BH) of the stars within
    r_cut, the set taken at the given state."""
    r = x[1:] - x[0]
    w = v[1:] - v[0]
    R = torch.sqrt((r * r).sum(1))
    sel = (R < r_cut).detach()
    vr = (w * r).sum(1) / R
    return (vr[sel] ** 2).sum() / sel.sum()
