"""CPU restatement of the warm-start network (d3-force-wasm_amd/csrc/gnn.hip)
for checking its GPU forward and backward passes.

Test infrastructure only: imported by tests/, never by the product. The
network has no reference counterpart (the reference has no learned warm
start), so it is pinned two ways: this restatement's analytic backward
against central finite differences in FP64 (tests/test_oracle_gnn.py), and
the GPU kernels against this restatement with BF16 rounding emulated at the
GPU's rounding points (bf16=True).

Parameter layout (FP32 on the GPU): W0[16][64] (rows >= 8 zero), b0[64],
then for each of 4 layers Ws[64][64], Wn[64][64], b[64], then Wo[64][16]
(columns >= 2 zero), bo[16].
"""

import numpy as np

F_IN, K_IN, HID, LAYERS, OUT, OUT_PAD = 8, 16, 64, 4, 2, 16
OFF_W0 = 0
OFF_B0 = OFF_W0 + K_IN * HID
OFF_L = OFF_B0 + HID
L_STRIDE = 2 * HID * HID + HID
OFF_WO = OFF_L + LAYERS * L_STRIDE
OFF_BO = OFF_WO + HID * OUT_PAD
N_PARAMS = OFF_BO + OUT_PAD


def bf16(a):
    """Round to BF16 (nearest, ties to even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w0 = p[OFF_W0:OFF_B0].reshape(K_IN, HID)
    b0 = p[OFF_B0:OFF_L]
    layers = []
    for l in range(LAYERS):
        o = OFF_L + l * L_STRIDE
        layers.append((p[o:o + HID * HID].reshape(HID, HID), p[o + HID * HID:o + 2 * HID * HID].reshape(HID, HID),
                       p[o + 2 * HID * HID:o + L_STRIDE]))
    wo = p[OFF_WO:OFF_BO].reshape(HID, OUT_PAD)
    bo = p[OFF_BO:OFF_BO + OUT_PAD]
    return w0, b0, layers, wo, bo


def mean_rows(h, row, col, transpose=False):
    """mean over CSR neighbours; transpose: sum over neighbours j of h[j] / deg(j)."""
    n = len(row) - 1
    deg = np.diff(row.astype(np.int64)).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    out = np.zeros((n, h.shape[1]))
    if transpose:
        np.add.at(out, src, h[col] / deg[col][:, None])
    else:
        np.add.at(out, src, h[col])
        out /= np.maximum(deg, 1)[:, None]
    return out


def forward(p, x, row, col, bf=False):
    """Positions (n x 2) and the activations the backward pass needs."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    w0, b0, layers, wo, bo = unpack(p)
    xp = np.zeros((x.shape[0], K_IN))
    xp[:, :F_IN] = x
    xp = r(xp)
    z = [xp @ r(w0) + b0]
    h = [None, r(np.maximum(z[0], 0))]
    agg = []
    for l, (ws, wn, b) in enumerate(layers):
        a = r(mean_rows(h[l + 1], row, col))
        agg.append(a)
        z.append(h[l + 1] @ r(ws) + a @ r(wn) + b)
        h.append(r(np.maximum(z[-1], 0)))
    pos = h[LAYERS + 1] @ r(wo) + bo
    return pos[:, :OUT], (xp, z, h, agg)


def backward(p, x, row, col, dp, bf=False):
    """Gradient of sum(dp * positions) with respect to the parameters."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    w0, b0, layers, wo, bo = unpack(p)
    _, (xp, z, h, agg) = forward(p, x, row, col, bf)
    g = np.zeros(N_PARAMS)
    dpp = np.zeros((x.shape[0], OUT_PAD))
    dpp[:, :OUT] = dp
    dpb = r(dpp)
    g[OFF_WO:OFF_BO] = (h[LAYERS + 1].T @ dpb).ravel()
    g[OFF_BO:OFF_BO + OUT_PAD] = dpp.sum(axis=0)
    dh = dpb @ r(wo).T
    for l in range(LAYERS - 1, -1, -1):
        ws, wn, b = layers[l]
        o = OFF_L + l * L_STRIDE
        dzf = np.where(z[l + 1] > 0, dh, 0.0)
        dz = r(dzf)
        g[o:o + HID * HID] = (h[l + 1].T @ dz).ravel()
        g[o + HID * HID:o + 2 * HID * HID] = (agg[l].T @ dz).ravel()
        g[o + 2 * HID * HID:o + L_STRIDE] = dzf.sum(axis=0)
        t = r(mean_rows(r(dz @ r(wn).T), row, col, transpose=True))
        dh = dz @ r(ws).T + t
    dzf = np.where(z[0] > 0, dh, 0.0)
    dz = r(dzf)
    gw0 = xp.T @ dz
    gw0[F_IN:] = 0
    g[OFF_W0:OFF_B0] = gw0.ravel()
    g[OFF_B0:OFF_L] = dzf.sum(axis=0)
    g[OFF_WO:OFF_BO].reshape(HID, OUT_PAD)[:, OUT:] = 0
    g[OFF_BO + OUT:OFF_BO + OUT_PAD] = 0
    return g
