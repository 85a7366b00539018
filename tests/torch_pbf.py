"""An independent dense float64 PBF step in torch, used as the reference gradient.

Same semantics as the reference step (predict and clamp, grid candidates from the
predicted positions, I iterations of lambda / delta, clamp, velocity, XSPH):
- the candidate pairs are the 27-cell neighbourhood of the grid built on the
  predicted positions;
- the in-range test r^2 <= h^2 is redone at each stage;
- lambda takes the first maxNeighbors in-range neighbours in walk order, which is
  the order of the stable sort by row-major cell key (rows of the 27-cell walk are
  x-lines of cells visited in increasing key order).
It is written with dense N x N tensors, so it is meant for small N.
torch.autograd differentiates it; its own gradients are checked against float64
central differences in tests.
"""
import torch


def pbf_step(X, V, sp, c, box_min, box_max):
    dt_ = X.dtype
    h = float(sp.kernel.h)
    h2 = h * h
    poly6, spiky = float(sp.kernel.poly6), float(sp.kernel.spiky)
    m, rho0, dt = float(sp.particleMass), float(sp.restDensity), float(sp.dt)
    pbf = sp.pbf
    g = torch.tensor(list(sp.gravity), dtype=dt_)
    lo = torch.tensor(box_min, dtype=dt_)
    hi = torch.tensor(box_max, dtype=dt_)
    n = X.shape[0]
    eye = torch.eye(n, dtype=torch.bool)

    P = X + (V + g * dt) * dt
    P = torch.minimum(torch.maximum(P, lo), hi)
    cs = float(sp.grid.cellSize)
    dims = torch.tensor(list(sp.grid.dim))
    cell = torch.minimum(torch.maximum(torch.floor((P.detach() - lo) / cs).long(), torch.zeros(3, dtype=torch.long)),
                         dims - 1)
    cand = ((cell[:, None, :] - cell[None, :, :]).abs() <= 1).all(-1) & ~eye
    key = (cell[:, 2] * dims[1] + cell[:, 1]) * dims[0] + cell[:, 0]
    order = torch.argsort(key * n + torch.arange(n), stable=True)  # slot -> particle
    rank = torch.empty_like(order)
    rank[order] = torch.arange(n)                                     # particle -> slot
    cap = int(sp.maxNeighbors)

    def capped(nb):
        if cap <= 0:
            return nb
        s_nb = nb[:, order]
        keep = s_nb & (s_nb.long().cumsum(1) <= cap)
        return keep[:, rank]

    def pairs(P):
        d = P[:, None, :] - P[None, :, :]
        r2 = (d * d).sum(-1)
        nb = cand & (r2 <= h2)
        # coincident neighbours (two particles clamped into the same box corner) have
        # r = 0 and a zero gradient vector; the square root is taken of a safe value so
        # its derivative at 0 (infinite) never meets a zero in the backward pass
        pos = nb & (r2 > 0)
        r = torch.sqrt(torch.where(pos, r2, torch.ones_like(r2))) * pos.to(r2.dtype)
        return d, r2, r, nb

    dq = pbf.scorr_dq_h * h
    wq = poly6 * (h2 - dq * dq) ** 3
    maxDisp = pbf.disp_clamp_max_h * h
    for _ in range(sp.solverIters):
        d, r2, r, nb = pairs(P)
        nbf = nb.to(dt_)
        nc = capped(nb)
        ncf = nc.to(dt_)
        W = poly6 * torch.clamp(h2 - r2, min=0.0) ** 3 * nbf
        rho = m * (W * ncf).sum(1)
        C = rho / rho0 - 1.0
        rs = torch.clamp(r, min=pbf.grad_r_eps)
        coeff = (-3.0 * spiky * (h - r) ** 2 / rs) * nbf
        gr = coeff[..., None] * d
        grc = gr * ncf[..., None]
        S1 = (grc * grc).sum(-1).sum(1)
        G = grc.sum(1)
        scale = m / rho0
        den = scale * scale * (S1 + (G * G).sum(1)) + pbf.lambda_denom_eps
        if pbf.compliance > 0 and dt > 0:
            den = den + pbf.compliance / (dt * dt)
        den = torch.clamp(den, min=pbf.lambda_denom_eps)
        lam = -C / den
        gate = (nc.sum(1) >= 2) & ((S1 + (G * G).sum(1)) >= 1e-12)
        lam = torch.where(gate, lam, torch.zeros_like(lam))
        if pbf.enable_lambda_clamp:
            lam = torch.clamp(lam, -pbf.lambda_max_abs, pbf.lambda_max_abs)
        scorr = torch.zeros_like(W)
        if pbf.scorr_enable:
            scorr = -pbf.scorr_k * (W / wq) ** pbf.scorr_n
            if pbf.scorr_min < 0:
                scorr = torch.clamp(scorr, min=pbf.scorr_min)
        lij = (lam[:, None] + lam[None, :] + scorr) * nbf
        D = (lij[..., None] * gr).sum(1) / rho0
        if pbf.enable_relax:
            D = D * pbf.relax_omega
        if pbf.enable_disp_clamp:
            l2 = (D * D).sum(1, keepdim=True)
            over = l2 > maxDisp * maxDisp
            # the unclamped branch's backward must not see maxDisp / 0 (an isolated
            # particle has D = 0): the length is taken of a safe value
            ln = torch.sqrt(torch.where(over, l2, torch.ones_like(l2)))
            D = torch.where(over, D * (maxDisp / ln), D)
        P = P + D
    P = torch.minimum(torch.maximum(P, lo), hi)
    Vn = (P - X) / dt
    d, r2, r, nb = pairs(P)
    W = poly6 * torch.clamp(h2 - r2, min=0.0) ** 3 * nb.to(dt_)
    dv = (W[..., None] * (Vn[None, :, :] - Vn[:, None, :])).sum(1)
    gate = torch.ones(n, dtype=dt_)
    if pbf.xsph_gate_enable:
        n_min = max(0, pbf.xsph_n_min)
        n_max = max(n_min + 1, pbf.xsph_n_max)
        gate = torch.clamp((nb.sum(1).to(dt_) - n_min) / (n_max - n_min), 0.0, 1.0)
    Vout = Vn + c * (gate * (m / rho0))[:, None] * dv
    return P, Vout


def rollout(X, V, sp, c, steps, box_min, box_max):
    for _ in range(steps):
        X, V = pbf_step(X, V, sp, c, box_min, box_max)
    return X, V
