"""FP64 reference for the differentiable rollout (test infrastructure).

An independent float64 torch implementation of the smooth gradient step — policies
{SocialForcesAvoidance + GoalReachingForce} with the gradient method, contact forces,
the acceleration and speed clamps, integration and the toric wrap — whose gradients
come from torch autograd. Neighbour lists are recomputed every step by brute force in
float64 with the reference's rules (toric axis queries, self excluded, canonical order,
cap); like the GPU adjoint, autograd holds the lists and every branch fixed. Its forward
step is checked against the FP64 oracle in tests/test_adjoint_cpu.py.

Formulas: SocialForcesAvoidance.cpp:26-110, GoalReachingForce.cpp:26-29,
ForceBasedFunction.cpp:32-36,49-52, policy.cpp:43-89,107-141, agent.cpp:52-59,71-93,
worldToric.cpp:61-108.
"""
import math

import numpy as np
import torch

F64 = torch.float64


def neighbor_lists(pos, world_type, W, H, rng, cap):
    """(N, K) neighbour indices, (N, K, 2) phantom offsets, (N, K) mask; canonical order."""
    p = pos.detach().cpu().numpy()
    n = len(p)
    r2 = float(np.float32(rng) * np.float32(rng))
    disps = [(0.0, 0.0)]
    lists = []
    maxk = 0
    for i in range(n):
        segs = [(0.0, 0.0)]
        if world_type == 1:
            if p[i, 0] - rng < -0.5 * W: segs.append((W, 0.0))
            if p[i, 0] + rng > 0.5 * W: segs.append((-W, 0.0))
            if p[i, 1] - rng < -0.5 * H: segs.append((0.0, H))
            if p[i, 1] + rng > 0.5 * H: segs.append((0.0, -H))
        entries = []
        for dx, dy in segs:
            q = p[i] + np.array([dx, dy])
            d2 = ((q[None, :] - p) ** 2).sum(1)
            idx = np.where(d2 < r2)[0]
            idx = idx[idx != i]
            order = np.lexsort((idx, d2[idx]))
            entries += [(int(j), -dx, -dy) for j in idx[order]]
        if cap > 0:
            entries = entries[:cap]
        lists.append(entries)
        maxk = max(maxk, len(entries))
    K = max(maxk, 1)
    idx = np.zeros((n, K), np.int64)
    off = np.zeros((n, K, 2))
    mask = np.zeros((n, K), bool)
    for i, e in enumerate(lists):
        for k, (j, ox, oy) in enumerate(e):
            idx[i, k] = j
            off[i, k] = (ox, oy)
            mask[i, k] = True
    return torch.from_numpy(idx), torch.from_numpy(off), torch.from_numpy(mask)


def _norm(v):
    """|v| with a finite gradient at v = 0 (where torch.where masks are applied)."""
    s = (v * v).sum(-1)
    ok = s > 0
    return torch.where(ok, torch.sqrt(torch.where(ok, s, torch.ones_like(s))), torch.zeros_like(s))


def _safe_normalize(v):
    m = _norm(v)
    ok = m > 0
    return torch.where(ok[..., None], v / torch.where(ok, m, torch.ones_like(m))[..., None], torch.zeros_like(v)), m


def _clamp(v, L):
    m = _norm(v)
    big = m > L
    scale = torch.where(big, L / torch.where(big, m, torch.ones_like(m)), torch.ones_like(m))
    return v * scale[..., None]


def step(pos, vel, A, B, c):
    """One step. c: dict with goal, radius, pref_speed, max_speed, max_accel (tensors),
    dt, relax, mass_sf, sdt, rng, contact, world_type, W, H, cap, coeff_sf, coeff_gr."""
    n = pos.shape[0]
    dt = c["dt"]
    idx, off, mask = neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])
    # preferred velocity (agent.cpp:52-59, 99-102)
    d = c["goal"] - pos
    reached = (d * d).sum(-1) <= c["radius"] ** 2
    dn, _ = _safe_normalize(d)
    pref = torch.where(reached[:, None], torch.zeros_like(d), dn * c["pref_speed"][:, None])
    pj = pos[idx] + off                                   # phantom positions (N, K, 2)
    vj = vel[idx]
    prefj = pref[idx]
    rj = c["radius"][idx]
    ri = c["radius"][:, None]
    Rv = pos[:, None, :] - pj
    dsq = (Rv * Rv).sum(-1)
    inrange = mask & (dsq <= c["rng"] ** 2)
    # SocialForcesAvoidance pair forces
    magR = _norm(Rv)
    act = inrange & (magR > ri + rj)
    vb = _norm(vj)
    Eb, _ = _safe_normalize(prefj)
    V = (vb * c["sdt"])[..., None] * Eb
    magV = _norm(V)
    R2 = Rv - V
    magR2 = _norm(R2)
    bSq2 = (magR + magR2) ** 2 - (vb * c["sdt"]) ** 2
    act = act & (bSq2 > 0)
    b = 0.5 * torch.sqrt(torch.where(act, bSq2, torch.ones_like(bSq2)))
    act = act & (b >= 0.001)
    safe = lambda x: torch.where(act, x, torch.ones_like(x))
    cosang = torch.where(magV > 0, (Rv * V).sum(-1) / safe(magR * torch.where(magV > 0, magV, torch.ones_like(magV))),
                         torch.zeros_like(magV))
    d_magR2 = 1.0 / (2.0 * safe(magR2)) * (2.0 * magR - 2.0 * magV * cosang)
    d_bSq2 = 2.0 * (magR + magR2) * (1.0 + d_magR2)
    d_b = (1.0 / (8.0 * safe(b))) * d_bSq2
    d_V = -A / B * torch.exp(-b / B) * d_b
    # view factor: angle(v_i, -R) < 100 degrees (piecewise constant)
    vi = vel[:, None, :].expand_as(Rv)
    lv = _norm(vi) * magR
    cosv = torch.where(lv > 0, (vi * (-Rv)).sum(-1) / torch.where(lv > 0, lv, torch.ones_like(lv)), torch.ones_like(lv))
    ang = torch.where(lv > 0, torch.arccos(torch.clamp(cosv.detach(), -1.0, 1.0)), torch.zeros_like(lv))
    scale = torch.where(ang < 100.0 * math.pi / 180.0, 1.0, 0.5)
    F = torch.where(act[..., None], -Rv * (d_V * scale)[..., None], torch.zeros_like(Rv)).sum(1)
    F_gr = (pref - vel) / max(c["relax"], dt)
    acc = c["coeff_sf"] * F / c["mass_sf"] + c["coeff_gr"] * F_gr
    acc = _clamp(acc, c["max_accel"])
    nxt = _clamp(vel + acc * dt, c["max_speed"])
    nxt = torch.where((pref == 0).all(-1)[:, None], torch.zeros_like(nxt), nxt)
    # contact forces over every listed neighbour
    diff = Rv
    dm = _norm(diff)
    inter = ri + rj - dm
    cact = mask & (inter > 0) & (dm > 0)
    un = diff / torch.where(cact, dm, torch.ones_like(dm))[..., None]
    contact = torch.where(cact[..., None], un * (c["contact"] * inter)[..., None], torch.zeros_like(un)).sum(1)
    v_new = nxt + contact / 80.0 * dt
    p_new = pos + v_new * dt
    if c["world_type"] == 1:
        W, H = c["W"], c["H"]
        x = torch.where(p_new[:, 0] > 0.5 * W, p_new[:, 0] - W,
                        torch.where(p_new[:, 0] < -0.5 * W, p_new[:, 0] + W, p_new[:, 0]))
        y = torch.where(p_new[:, 1] > 0.5 * H, p_new[:, 1] - H,
                        torch.where(p_new[:, 1] < -0.5 * H, p_new[:, 1] + H, p_new[:, 1]))
        p_new = torch.stack([x, y], -1)
    return p_new, v_new


def consts_from(sc):
    """Constants of step() from a calibration Scenario (one SocialForces policy)."""
    a = sc.agents
    pol = sc.policies[0]
    c_sf = [i for i in range(pol["n_cost"]) if pol["cost_kind"][i] == 2][0]
    c_gr = [i for i in range(pol["n_cost"]) if pol["cost_kind"][i] == 0][0]
    t = lambda x: torch.tensor(np.asarray(x, np.float64), dtype=F64)
    return {"goal": t(a["goal"]), "radius": t(a["radius"]), "pref_speed": t(a["pref_speed"]),
            "max_speed": t(a["max_speed"]), "max_accel": t(a["max_accel"]), "dt": float(np.float32(sc.dt)),
            "relax": float(np.float32(pol["relaxation_time"])), "mass_sf": float(pol["cost_params"][c_sf][0]),
            "sdt": float(pol["cost_params"][c_sf][1]), "rng": float(np.float32(pol["cost_range"][c_sf])),
            "contact": float(pol["contact_force_scale"]), "world_type": sc.world_type,
            "W": float(np.float32(sc.width)), "H": float(np.float32(sc.height)), "cap": sc.max_neighbors,
            "coeff_sf": float(pol["cost_coeff"][c_sf]), "coeff_gr": float(pol["cost_coeff"][c_gr])}


def minimg(d, c):
    if c["world_type"] != 1:
        return d
    W, H = c["W"], c["H"]
    x = torch.where(d[:, 0] > 0.5 * W, d[:, 0] - W, torch.where(d[:, 0] < -0.5 * W, d[:, 0] + W, d[:, 0]))
    y = torch.where(d[:, 1] > 0.5 * H, d[:, 1] - H, torch.where(d[:, 1] < -0.5 * H, d[:, 1] + H, d[:, 1]))
    return torch.stack([x, y], -1)


def rollout_loss(pos0, v0, A, B, c, targets):
    pos, vel = pos0, v0
    T = targets.shape[0]
    n = pos0.shape[0]
    loss = torch.zeros((), dtype=F64)
    for t in range(T):
        pos, vel = step(pos, vel, A, B, c)
        dd = minimg(pos - targets[t], c)
        loss = loss + (dd * dd).sum()
    return loss / (T * n)
