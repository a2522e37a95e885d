"""Test infrastructure only: a numpy restatement of the learned preconditioner's
graph network (forward, loss and backward), the checker of the HIP kernels in
xhdfe-xfe_amd/csrc/lp_gnn.hip.  Nothing in the product imports it.

The reference has no learned component: its preconditioner for the FE normal
equations is the randomized approximate Cholesky of src/schwarz_demean.cpp:100-133
inside the block PCG of :58-415.  The network here is this build's own design:

Graph (the worker-firm mobility graph): firms, movers (workers with two or more
firm segments), one edge per (mover, firm).  Message passing, width H = 64:

    hf = relu(xf Wf0 + bf0),  hw = relu(xw Ww0 + bw0)              (inputs: 8 features each)
    for l = 1..3:
        hw = relu([hw | mean_{f in N(w)} hf] Wwl + bwl)             (Wwl: 128 x 64)
        hf = relu([hf | mean_{w in N256(f)} hw] Wfl + bfl)          (Wfl: 128 x 64)

N256(f): the first 256 movers of f's list (ascending ids), a fixed-size
neighbour sample that keeps one firm's mean from serializing a launch.
    u_f  = hf . wa + ba                                             (per-level score)
    t_fg = sum_k hf_k hg_k we_k + be + ws q_fg                      (pair score; q_fg the pair's feature)

Loss: mean over firms with a target of (u_f - u*_f)^2, plus the mean binary
cross-entropy of sigmoid(t_fg) against y*_fg over the firm pairs.

With bf16=True every matrix product rounds its two operands to bfloat16
(round to nearest even), as the device's BF16 MFMA instructions take them, and
accumulates exactly (float64): the device differs only by its FP32
accumulation order.  With bf16=False the function is smooth in the parameters,
which is what the finite-difference check of the gradients needs.
Parity: gradients pinned by central finite differences (tests/test_oracle_cpu.py).
"""
from __future__ import annotations

import numpy as np

H = 64
NF_IN = 8   # firm input features
NW_IN = 8   # mover input features
LAYERS = 3

# parameter layout (name, shape), packed in this order into one float32 vector
LAYOUT = ([("Wf0", (NF_IN, H)), ("bf0", (H,)), ("Ww0", (NW_IN, H)), ("bw0", (H,))]
          + [item for l in range(1, LAYERS + 1)
             for item in ((f"Ww{l}", (2 * H, H)), (f"bw{l}", (H,)), (f"Wf{l}", (2 * H, H)), (f"bf{l}", (H,)))]
          + [("wa", (H,)), ("ba", (1,)), ("we", (H,)), ("be", (1,)), ("ws", (1,))])
NPARAM = int(sum(np.prod(s) for _, s in LAYOUT))


def unpack(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def init_params(seed=0):
    """He-normal weights, zero biases (float32 vector of NPARAM)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in LAYOUT:
        if len(shape) == 2:
            d[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])
        elif name in ("wa", "we"):
            d[name] = rng.standard_normal(shape) * np.sqrt(1.0 / H)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def bf16(a):
    """Round to bfloat16 (nearest, ties to even) and back to float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, rb):
    return (bf16(a) @ bf16(b)) if rb else (a @ b)


FIRM_CAP = 256  # a firm's mean over movers takes the first FIRM_CAP of its list


def _capped(ptr, cap):
    """Rows, positions kept and capped degrees of a CSR list."""
    deg = np.diff(ptr)
    n = ptr.size - 1
    rows = np.repeat(np.arange(n), deg)
    pos = np.arange(rows.size) - np.repeat(ptr[:-1], deg)
    keep = pos < cap
    return rows, keep, np.minimum(deg, cap)


def _mean_gather(ptr, idx, src, cap=None):
    """out[v] = mean of src over the first cap entries of idx[ptr[v]:ptr[v+1]]
    (all without cap; 0 for no neighbours)."""
    n = ptr.size - 1
    rows, keep, d = _capped(np.asarray(ptr, np.int64), cap if cap is not None else np.iinfo(np.int64).max)
    out = np.zeros((n, src.shape[1]))
    np.add.at(out, rows[keep], src[np.asarray(idx)[keep]])
    return out / np.maximum(d, 1)[:, None]


def _mean_scatter(ptr, idx, g, nsrc, cap=None):
    """Transpose of _mean_gather: d src from d out."""
    rows, keep, d = _capped(np.asarray(ptr, np.int64), cap if cap is not None else np.iinfo(np.int64).max)
    out = np.zeros((nsrc, g.shape[1]))
    r = rows[keep]
    np.add.at(out, np.asarray(idx)[keep], g[r] / np.maximum(d, 1)[r, None])
    return out


def forward_backward(theta, g, bf16_products=True, want_grad=True):
    """Loss, gradient (float64 vector of NPARAM) and outputs (u, t).

    g: dict with xf (F x 8), xw (Wm x 8), f_ptr/f_idx (firm -> mover edges,
    CSR over firms, entries mover ids), w_ptr/w_idx (mover -> firm edges), pf,
    pg (firm pairs), u_tgt (F), u_mask (F, 1 where the firm has a target),
    y_tgt (pairs)."""
    rb = bf16_products
    P = unpack(theta)
    xf, xw = np.asarray(g["xf"], np.float64), np.asarray(g["xw"], np.float64)
    F, Wm = xf.shape[0], xw.shape[0]
    cache = []
    zf = _mm(xf, P["Wf0"], rb) + P["bf0"]
    hf = np.maximum(zf, 0.0)
    zw = _mm(xw, P["Ww0"], rb) + P["bw0"]
    hw = np.maximum(zw, 0.0)
    for l in range(1, LAYERS + 1):
        aw = _mean_gather(g["w_ptr"], g["w_idx"], hf)
        inw = np.hstack([hw, aw])
        zw = _mm(inw, P[f"Ww{l}"], rb) + P[f"bw{l}"]
        hw_new = np.maximum(zw, 0.0)
        af = _mean_gather(g["f_ptr"], g["f_idx"], hw_new, FIRM_CAP)
        inf = np.hstack([hf, af])
        zf = _mm(inf, P[f"Wf{l}"], rb) + P[f"bf{l}"]
        hf_new = np.maximum(zf, 0.0)
        cache.append((inw, zw, inf, zf))
        hw, hf = hw_new, hf_new
    u = hf @ P["wa"] + P["ba"][0]
    pf, pg = np.asarray(g["pf"]), np.asarray(g["pg"])
    prod = hf[pf] * hf[pg]
    pq = np.asarray(g.get("pq", np.zeros(pf.size)), np.float64)
    t = prod @ P["we"] + P["be"][0] + P["ws"][0] * pq
    m = np.asarray(g["u_mask"], np.float64)
    nm = max(m.sum(), 1.0)
    du = 2.0 * m * (u - g["u_tgt"]) / nm
    Lu = float((m * (u - g["u_tgt"]) ** 2).sum() / nm)
    y = np.asarray(g["y_tgt"], np.float64)
    npair = max(t.size, 1)
    # BCE with logits, stable: max(t,0) - t y + log(1 + exp(-|t|))
    Le = float((np.maximum(t, 0.0) - t * y + np.log1p(np.exp(-np.abs(t)))).sum() / npair) if t.size else 0.0
    dt = (1.0 / (1.0 + np.exp(-t)) - y) / npair
    loss = Lu + Le
    out = dict(u=u, t=t, loss_u=Lu, loss_e=Le)
    if not want_grad:
        return loss, None, out
    G = {name: np.zeros(shape) for name, shape in LAYOUT}
    G["wa"] = hf.T @ du
    G["ba"] = np.array([du.sum()])
    G["we"] = prod.T @ dt
    G["be"] = np.array([dt.sum()])
    G["ws"] = np.array([(dt * pq).sum()])
    dhf = np.outer(du, P["wa"])
    dprod = np.outer(dt, P["we"])
    np.add.at(dhf, pf, dprod * hf[pg])
    np.add.at(dhf, pg, dprod * hf[pf])
    dhw = np.zeros((Wm, H))
    for l in range(LAYERS, 0, -1):
        inw, zw, inf, zf = cache[l - 1]
        dzf = dhf * (zf > 0)
        G[f"Wf{l}"] = _mm(inf.T, dzf, rb)
        G[f"bf{l}"] = dzf.sum(axis=0)
        dinf = _mm(dzf, P[f"Wf{l}"].T, rb)
        dhf_prev = dinf[:, :H]
        dhw_new = dhw + _mean_scatter(g["f_ptr"], g["f_idx"], dinf[:, H:], Wm, FIRM_CAP)
        dzw = dhw_new * (zw > 0)
        G[f"Ww{l}"] = _mm(inw.T, dzw, rb)
        G[f"bw{l}"] = dzw.sum(axis=0)
        dinw = _mm(dzw, P[f"Ww{l}"].T, rb)
        dhw = dinw[:, :H]
        dhf = dhf_prev + _mean_scatter(g["w_ptr"], g["w_idx"], dinw[:, H:], F)
    zf0 = _mm(xf, P["Wf0"], rb) + P["bf0"]
    dz = dhf * (zf0 > 0)
    G["Wf0"] = _mm(xf.T, dz, rb)
    G["bf0"] = dz.sum(axis=0)
    zw0 = _mm(xw, P["Ww0"], rb) + P["bw0"]
    dz = dhw * (zw0 > 0)
    G["Ww0"] = _mm(xw.T, dz, rb)
    G["bw0"] = dz.sum(axis=0)
    return loss, pack(G), out


def random_graph(seed, F=40, Wm=60, deg=(2, 4)):
    """A small random mobility graph with features and targets (tests)."""
    rng = np.random.default_rng(seed)
    edges = []
    for w in range(Wm):
        k = int(rng.integers(deg[0], deg[1] + 1))
        edges += [(w, int(f)) for f in rng.choice(F, size=k, replace=False)]
    e = np.array(edges)
    g = csr_pair(F, Wm, e[:, 0], e[:, 1])
    g["xf"] = rng.standard_normal((F, NF_IN)).astype(np.float32)
    g["xw"] = rng.standard_normal((Wm, NW_IN)).astype(np.float32)
    g["u_tgt"] = rng.standard_normal(F)
    g["u_mask"] = (rng.random(F) < 0.8).astype(np.float64)
    g["y_tgt"] = rng.random(g["pf"].size)
    g["pq"] = rng.random(g["pf"].size).astype(np.float32)
    return g


def csr_pair(F, Wm, ew, ef):
    """Both CSR directions of the mover-firm edges and the firm pairs of
    consecutive edges of one mover."""
    o = np.lexsort((ef, ew))
    ew, ef = ew[o], ef[o]
    w_ptr = np.zeros(Wm + 1, dtype=np.int32)
    np.cumsum(np.bincount(ew, minlength=Wm), out=w_ptr[1:])
    o2 = np.argsort(ef, kind="stable")
    f_ptr = np.zeros(F + 1, dtype=np.int32)
    np.cumsum(np.bincount(ef, minlength=F), out=f_ptr[1:])
    same = ew[1:] == ew[:-1]
    return dict(w_ptr=w_ptr, w_idx=ef.astype(np.int32), f_ptr=f_ptr, f_idx=ew[o2].astype(np.int32),
                pf=ef[:-1][same].astype(np.int32), pg=ef[1:][same].astype(np.int32))
