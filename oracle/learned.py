"""FP64 numpy oracle of the learned proposal (test infrastructure only).

No reference counterpart: the proposal network is new (a 3 -> 64 -> 64 -> 2 tanh MLP on
per-particle features), trained to maximise the expected squared jump distance of an exact
Metropolis-Hastings move.  This module restates the definitions the GPU kernels follow:

features of particle i at position z (others fixed), D_CAP = 8:
    f1 = |z| / R
    d_(1) <= ... <= d_(8) = the distances |z - z_j| (j != i) to the 8 nearest particles,
    each capped at D_CAP (fewer than 8 others: the missing ones count as D_CAP)
    f2 = d_(1),   f3 = mean_k d_(k)
widths: (sigma_r, sigma_t) = sigma0 * exp(ln(2.5) tanh(MLP(f))), within [0.4, 2.5] sigma0
move: z' = z + e^{i theta} (sigma_r n1 + i sigma_t n2),  theta = arg z
log q(z'|x) = -log(2 pi sigma_r sigma_t) - |n|^2 / 2
log q(z|x') = -log(2 pi s'_r s'_t) - ((D_r / s'_r)^2 + (D_t / s'_t)^2) / 2,
    D_r + i D_t = (z - z') e^{-i theta'}, widths s' from the features at z'
log alpha = m sum_j log(|z'-z_j|^2 / |z-z_j|^2) - (|z'|^2 - |z|^2) / 2 + log q(z|x') - log q(z'|x)
loss = -mean min(1, e^{log alpha}) |z' - z|^2
Parameters are packed as W1 (64x3), b1 (64), W2 (64x64), b2 (64), W3 (2x64), b3 (2): 4546.
"""
import numpy as np

D_CAP = 8.0
H = 64
NPARAM = H * 3 + H + H * H + H + 2 * H + 2
WSCALE = 0.91629073187415511   # ln 2.5


def unpack(p):
    o = 0
    W1 = p[o:o + H * 3].reshape(H, 3); o += H * 3
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + 2 * H].reshape(2, H); o += 2 * H
    b3 = p[o:o + 2]
    return W1, b1, W2, b2, W3, b3


def features(zs, i, z, R):
    """features and their gradient w.r.t. (Re z, Im z): f (3,), df (3, 2).
    d_(1) <= ... <= d_(8): distances from z to the 8 nearest particles j != i, each capped at
    D_CAP (missing neighbours count as D_CAP); f2 = d_(1), f3 = mean d_(k); a capped
    distance has zero gradient."""
    f = np.zeros(3)
    df = np.zeros((3, 2))
    r = abs(z)
    f[0] = r / R
    if r > 0:
        df[0] = [z.real / (r * R), z.imag / (r * R)]
    others = np.delete(np.asarray(zs), i)
    diff = z - others
    d = np.abs(diff)
    order = np.argsort(d, kind="stable")[:8]
    near = np.full(8, D_CAP)
    grads = np.zeros((8, 2))
    for q, j in enumerate(order):
        if d[j] < D_CAP:
            near[q] = d[j]
            if d[j] > 0:
                grads[q] = [diff[j].real / d[j], diff[j].imag / d[j]]
    f[1] = near[0]
    df[1] = grads[0]
    f[2] = near.mean()
    df[2] = grads.mean(0)
    return f, df


def mlp(p, f):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(W1 @ f + b1)
    h2 = np.tanh(W2 @ h1 + b2)
    o = W3 @ h2 + b3
    return h1, h2, o


def mlp_backward(p, f, h1, h2, do):
    """grads of sum(do * o) w.r.t. params (packed) and input f"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    gW3 = np.outer(do, h2)
    gb3 = do
    dh2 = W3.T @ do
    dz2 = dh2 * (1 - h2 ** 2)
    gW2 = np.outer(dz2, h1)
    gb2 = dz2
    dh1 = W2.T @ dz2
    dz1 = dh1 * (1 - h1 ** 2)
    gW1 = np.outer(dz1, f)
    gb1 = dz1
    df = W1.T @ dz1
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return g, df


def widths(o, sigma0):
    t = np.tanh(o)
    s = sigma0 * np.exp(WSCALE * t)
    ds_do = s * WSCALE * (1 - t ** 2)
    return s, ds_do


def sample_loss_grad(p, zs, i, n, m, R, sigma0):
    """loss contribution -A J of one (configuration, particle, noise) triple and its gradient"""
    z = zs[i]
    th = np.angle(z) if z != 0 else 0.0
    c, s = np.cos(th), np.sin(th)
    f, _ = features(zs, i, z, R)
    h1, h2, o = mlp(p, f)
    sg, dsg = widths(o, sigma0)
    zp = z + (c + 1j * s) * (sg[0] * n[0] + 1j * sg[1] * n[1])
    fp, dfp = features(zs, i, zp, R)
    h1p, h2p, op = mlp(p, fp)
    sgp, dsgp = widths(op, sigma0)
    others = np.delete(zs, i)
    E = m * np.sum(np.log(np.abs(zp - others) ** 2 / np.abs(z - others) ** 2)) - 0.5 * (abs(zp) ** 2 - abs(z) ** 2)
    rp = abs(zp)
    cp, sp = (zp.real / rp, zp.imag / rp) if rp > 0 else (1.0, 0.0)
    w = z - zp
    Dr = w.real * cp + w.imag * sp
    Dt = -w.real * sp + w.imag * cp
    logq_rev = -np.log(2 * np.pi * sgp[0] * sgp[1]) - 0.5 * ((Dr / sgp[0]) ** 2 + (Dt / sgp[1]) ** 2)
    logq_fwd = -np.log(2 * np.pi * sg[0] * sg[1]) - 0.5 * (n[0] ** 2 + n[1] ** 2)
    la = E + logq_rev - logq_fwd
    A = min(1.0, np.exp(min(la, 0.0)))
    J = (sg[0] * n[0]) ** 2 + (sg[1] * n[1]) ** 2
    loss = -A * J
    # backward
    dla = -J * A if la < 0 else 0.0
    dJ = -A
    # through sigma' (reverse widths)
    dsgp_ = dla * np.array([-1 / sgp[0] + Dr ** 2 / sgp[0] ** 3, -1 / sgp[1] + Dt ** 2 / sgp[1] ** 3])
    dop = dsgp_ * dsgp
    g2, dfp_in = mlp_backward(p, fp, h1p, h2p, dop)
    # d/dz' (x', y'): energy, reverse density via D, features at z'
    dE = np.array([np.sum(2 * m * (zp - others).real / np.abs(zp - others) ** 2) - zp.real,
                   np.sum(2 * m * (zp - others).imag / np.abs(zp - others) ** 2) - zp.imag])
    a2, b2 = sgp[0] ** 2, sgp[1] ** 2
    dQ_dDr, dQ_dDt = Dr / a2, Dt / b2            # Q = (Dr^2/a2 + Dt^2/b2)/2, logq_rev = ... - Q
    if rp > 0:
        dDr = np.array([-cp - sp * Dt / rp, -sp + cp * Dt / rp])
        dDt = np.array([sp + sp * Dr / rp, -cp - cp * Dr / rp])
    else:
        dDr = np.array([-1.0, 0.0])
        dDt = np.array([0.0, -1.0])
    dzp = dla * (dE - (dQ_dDr * dDr + dQ_dDt * dDt)) + dfp.T @ dfp_in
    # d/dsigma (forward widths): logq_fwd, J, z'(sigma)
    dsg_ = dla * np.array([1 / sg[0], 1 / sg[1]]) + dJ * np.array([2 * sg[0] * n[0] ** 2, 2 * sg[1] * n[1] ** 2])
    dsg_ += np.array([dzp @ np.array([c, s]) * n[0], dzp @ np.array([-s, c]) * n[1]])
    do = dsg_ * dsg
    g1, _ = mlp_backward(p, f, h1, h2, do)
    return loss, g1 + g2, la


def batch_loss_grad(p, configs, idx, noise, m, R, sigma0):
    L = 0.0
    G = np.zeros_like(p)
    for b in range(len(idx)):
        l, g, _ = sample_loss_grad(p, configs[b], idx[b], noise[b], m, R, sigma0)
        L += l
        G += g
    return L / len(idx), G / len(idx)


def init_params(rng):
    """Xavier-uniform weights, zero biases (the GPU uses its own seeded Philox draw)."""
    p = np.zeros(NPARAM)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    W1[:] = rng.uniform(-1, 1, W1.shape) * np.sqrt(6 / (3 + H))
    W2[:] = rng.uniform(-1, 1, W2.shape) * np.sqrt(6 / (2 * H))
    W3[:] = rng.uniform(-1, 1, W3.shape) * np.sqrt(6 / (H + 2)) * 0.1
    return p
