"""TEST INFRASTRUCTURE ONLY: a torch (CPU, FP32) restatement of the device U-Net of hfk_net.hip, used to check the
device forward pass, loss and weight gradients.  BF16 rounding is emulated where the device stores BF16 (inputs,
weights, activations after every layer and transfer), so the comparison isolates indexing/adjoint mistakes from
rounding."""
import numpy as np
import torch
import torch.nn.functional as F

# (level, cin1, cin2, cout, relu): the layer table of hfk_net.hip
LAYERS = [(0, 16, 0, 16, True), (0, 16, 0, 16, True), (1, 16, 0, 32, True), (1, 32, 0, 32, True),
          (2, 32, 0, 64, True), (3, 64, 0, 64, True), (3, 64, 0, 64, True), (2, 64, 64, 64, True),
          (1, 32, 64, 32, True), (0, 16, 32, 16, True), (0, 16, 0, 16, False)]


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def unflatten(w):
    """Flat FP32 weights ([tap][co][ci] per layer) -> torch conv weights [co][ci][kz][ky][kx]."""
    out, off = [], 0
    for (_, c1, c2, co, _) in LAYERS:
        ci = c1 + c2
        n = 27 * co * ci
        blk = torch.tensor(np.asarray(w[off:off + n], dtype=np.float32)).reshape(3, 3, 3, co, ci)  # [kz][ky][kx]
        out.append(blk.permute(3, 4, 0, 1, 2).contiguous())
        off += n
    return out


def restrict(x):
    k1 = torch.tensor([0.25, 0.5, 0.25])
    k = torch.einsum("i,j,k->ijk", k1, k1, k1)
    C = x.shape[1]
    return F.conv3d(x, k.expand(C, 1, 3, 3, 3).contiguous(), stride=2, padding=1, groups=C)


def prolong(x):
    k1 = torch.tensor([0.5, 1.0, 0.5])
    k = torch.einsum("i,j,k->ijk", k1, k1, k1)
    C = x.shape[1]
    return F.conv_transpose3d(x, k.expand(C, 1, 3, 3, 3).contiguous(), stride=2, padding=1, groups=C)


def forward(Ws, x0):
    def conv(x, i):
        y = F.conv3d(x, bf(Ws[i]), padding=1)
        return bf(F.relu(y)) if LAYERS[i][4] else y

    a0 = conv(x0, 0)
    s0 = conv(a0, 1)
    r1 = bf(restrict(s0))
    a2 = conv(r1, 2)
    s1 = conv(a2, 3)
    r2 = bf(restrict(s1))
    s2 = conv(r2, 4)
    r3 = bf(restrict(s2))
    a5 = conv(r3, 5)
    a6 = conv(a5, 6)
    u2 = bf(prolong(a6))
    a7 = conv(torch.cat([s2, u2], 1), 7)
    u1 = bf(prolong(a7))
    a8 = conv(torch.cat([s1, u1], 1), 8)
    u0 = bf(prolong(a8))
    a9 = conv(torch.cat([s0, u0], 1), 9)
    return conv(a9, 10)


def loss_and_grad(w, node_dofs, N, h, r, e, coarsen=False):
    """Loss sum_free (out - t)^2 / sum t^2 with t = e h / s, s = RMS(r); and its gradient w.r.t. the weights.
    With ``coarsen`` the U-Net runs on every other vertex of the N^3 lattice: BF16 full-weighting restriction of
    the packed input, FP32 trilinear prolongation of the output."""
    Ws = [t.clone().requires_grad_(True) for t in unflatten(w)]
    s = np.sqrt(np.dot(r, r) / r.size)
    x0 = np.zeros((16, N ** 3), dtype=np.float32)
    t = np.zeros((3, N ** 3))
    mask = np.zeros((3, N ** 3), dtype=bool)
    for c in range(3):
        ids = node_dofs[:, c]
        f = ids > 0
        x0[c, f] = r[ids[f] - 1] / s
        t[c, f] = e[ids[f] - 1] * h / s
        mask[c, f] = True
    x0t = bf(torch.tensor(x0).reshape(1, 16, N, N, N))
    if coarsen:
        out = prolong(forward(Ws, bf(restrict(x0t)))[:, :3])[0].reshape(3, -1)
    else:
        out = forward(Ws, x0t)[0, :3].reshape(3, -1)
    tt = torch.tensor(t, dtype=torch.float32)
    mk = torch.tensor(mask)
    d = (out - tt)[mk]
    L = (d * d).sum() / float((t[mask] ** 2).sum())
    L.backward()
    g = torch.cat([W.grad.permute(2, 3, 4, 0, 1).reshape(-1) for W in Ws]).numpy()
    return float(L.detach()), g, out.detach().numpy()


def energy_loss_and_grad(w, node_dofs, N, h, A, r, coarsen=False):
    """The energy loss of the device training, L = (w'(A/h)w - 2 w'x) / n with x = r/s (s = RMS(r)) and w the
    network's output on the free dofs (n of them), and its gradient with respect to the weights.  A is the
    assembled coarse matrix (scipy sparse, free dofs)."""
    Ws = [t.clone().requires_grad_(True) for t in unflatten(w)]
    n = r.size
    s = np.sqrt(np.dot(r, r) / n)
    x0 = np.zeros((16, N ** 3), dtype=np.float32)
    for c in range(3):
        ids = node_dofs[:, c]
        f = ids > 0
        x0[c, f] = r[ids[f] - 1] / s
    x0t = bf(torch.tensor(x0).reshape(1, 16, N, N, N))
    if coarsen:
        out = prolong(forward(Ws, bf(restrict(x0t)))[:, :3])[0].reshape(3, -1)
    else:
        out = forward(Ws, x0t)[0, :3].reshape(3, -1)
    # gather the output at the free dofs: wv[id - 1] = out[c, vertex]
    idx_c, idx_v, idx_d = [], [], []
    for c in range(3):
        ids = node_dofs[:, c]
        f = np.nonzero(ids > 0)[0]
        idx_c.append(np.full(f.size, c))
        idx_v.append(f)
        idx_d.append(ids[f] - 1)
    idx_c, idx_v, idx_d = (np.concatenate(a) for a in (idx_c, idx_v, idx_d))
    order = np.argsort(idx_d)
    wv = out[torch.tensor(idx_c[order]), torch.tensor(idx_v[order])].to(torch.float64)
    Ac = A.tocoo()
    At = torch.sparse_coo_tensor(np.vstack([Ac.row, Ac.col]), Ac.data / h, (n, n)).coalesce()
    Aw = torch.sparse.mm(At, wv[:, None])[:, 0]
    x = torch.tensor(r / s)
    L = (wv @ Aw - 2.0 * (wv @ x)) / n
    L.backward()
    g = torch.cat([W.grad.permute(2, 3, 4, 0, 1).reshape(-1) for W in Ws]).numpy()
    return float(L.detach()), g
