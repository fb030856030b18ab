"""FP64 reference for the differentiable rollout (test infrastructure).

An independent float64 torch implementation of the smooth gradient step — policies
{SocialForcesAvoidance or PowerLaw, + GoalReachingForce} with the gradient method,
contact forces, the acceleration and speed clamps, integration and the toric wrap —
whose gradients come from torch autograd. Like the GPU adjoint, autograd holds the
neighbour lists and every branch of the step fixed.

Neighbour lists are either recomputed every step by brute force in float64 with the
reference's rules (toric axis queries, self excluded, canonical order, cap) or taken
from a recorded FP32 simulation (`lists_from_record`). Branch decisions (range filter,
overlap and cut-offs of the pair models, view factor, contact) are either taken on the
float64 values (branches="f64") or, for a state that an FP32 simulation produced,
evaluated in float32 with the FP32 kernels' operation order (branches="f32"), so that
the float64 values are differentiated on exactly the pieces the FP32 step used.
`replay_gradient` runs the float64 reverse sweep along a recorded FP32 trajectory: the
exact derivative, in float64, of the map the FP32 adjoint differentiates.

Its forward step is checked against the FP64 oracle, its gradients against central
differences, and its float32 branch decisions against its float64 ones (and against a
planted mistake) in tests/test_adjoint_cpu.py.

Formulas: SocialForcesAvoidance.cpp:26-110, PowerLaw.cpp:28-66,
GoalReachingForce.cpp:26-29, ForceBasedFunction.cpp:32-36,49-52,
ObjectInteractionForces.cpp:26-48, policy.cpp:43-89,107-141, agent.cpp:52-59,71-93,
worldToric.cpp:61-108.
"""
import math

import numpy as np
import torch

F64 = torch.float64
KIND_GR, KIND_PL, KIND_SF = 0, 1, 2


def view_cos_threshold():
    """The largest double c whose angle, as vector2D.h:177-190 computes it
    ((float)acos(c)), is not below (float)(100 deg): in view <=> cos > c."""
    half = np.float32(100.0 * 3.1415926535897 / 180.0)
    lo, hi = -1.0, 1.0
    while np.nextafter(lo, hi) < hi:
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            mid = float(np.nextafter(lo, hi))
        if np.float32(math.acos(mid)) >= half:
            lo = mid
        else:
            hi = mid
    return lo


VIEW_COS = view_cos_threshold()


def neighbor_lists(pos, world_type, W, H, rng, cap):
    """(N, K) neighbour indices, (N, K, 2) phantom offsets, (N, K) mask; canonical order."""
    p = pos.detach().cpu().numpy()
    n = len(p)
    r2 = float(np.float32(rng) * np.float32(rng))
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


def lists_from_record(cnt, ids, seg, W, H, row_of_id=None):
    """(idx, off, mask) from a recorded neighbour list (count, agent IDs and toric segments
    per agent; umans_dump_neighbors): segment 1 (+W query) places the neighbour at
    pos - (W, 0), 2 at pos + (W, 0), 3 at pos - (0, H), 4 at pos + (0, H)
    (worldToric.cpp:36,47). row_of_id maps IDs to rows (None: the ID is the row)."""
    cnt = np.asarray(cnt)
    K = max(int(cnt.max()) if len(cnt) else 0, 1)
    mask = np.arange(K)[None, :] < cnt[:, None]
    ids = np.where(mask, np.asarray(ids)[:, :K], 0)
    idx = ids if row_of_id is None else np.asarray(row_of_id)[ids]
    sg = np.where(mask, np.asarray(seg)[:, :K], 0)
    off = np.zeros(idx.shape + (2,))
    off[..., 0] = np.where(sg == 1, -W, np.where(sg == 2, W, 0.0))
    off[..., 1] = np.where(sg == 3, -H, np.where(sg == 4, H, 0.0))
    return torch.from_numpy(idx.astype(np.int64)), torch.from_numpy(off), torch.from_numpy(mask)


# ------------------------------------------------------------ branch decisions
def f32_decisions(pos, vel, idx, off, mask, c, p1):
    """The branch decisions of the FP32 step and adjoint kernels, evaluated in float32
    with their operation order from a float32 state: preferred velocity zero (goal
    reached), range filter, pair activity of the interaction model, view factor,
    contact. p1 = tau0 for PowerLaw (its cut-off)."""
    f = np.float32
    with np.errstate(all="ignore"):
        P = np.asarray(pos, np.float32)
        V = np.asarray(vel, np.float32)
        idx = np.asarray(idx)
        mask = np.asarray(mask)
        goal, rad, ps = c["goal32"], c["radius32"], c["pref_speed32"]
        # k_reorder: preferred velocity (agent.cpp:52-59, 99-102)
        tgx, tgy = goal[:, 0] - P[:, 0], goal[:, 1] - P[:, 1]
        sq = tgx * tgx + tgy * tgy
        reached = sq <= rad * rad
        m = np.sqrt(sq)
        nz = m > f(0)
        prx = np.where(nz, tgx / np.where(nz, m, f(1)), tgx) * ps
        pry = np.where(nz, tgy / np.where(nz, m, f(1)), tgy) * ps
        prx = np.where(reached, f(0), prx)
        pry = np.where(reached, f(0), pry)
        # the phantom neighbour: float(pos_j + offset) (NbList::get)
        o32 = np.asarray(off, np.float32)
        pjx = P[idx, 0] + o32[..., 0]
        pjy = P[idx, 1] + o32[..., 1]
        Rx = P[:, None, 0] - pjx
        Ry = P[:, None, 1] - pjy
        dsq = Rx * Rx + Ry * Ry
        rng = f(c["rng"])
        inrange = mask & (dsq <= rng * rng)
        ri = rad[:, None]
        rj = rad[idx]
        magR = np.sqrt(dsq)
        if c["kind"] == KIND_SF:
            vjx, vjy = V[idx, 0], V[idx, 1]
            vb = np.sqrt(vjx * vjx + vjy * vjy)
            pjx_, pjy_ = prx[idx], pry[idx]
            pm = np.sqrt(pjx_ * pjx_ + pjy_ * pjy_)
            pz = pm > f(0)
            ebx = np.where(pz, pjx_ / np.where(pz, pm, f(1)), pjx_)
            eby = np.where(pz, pjy_ / np.where(pz, pm, f(1)), pjy_)
            q = vb * f(c["sdt"])
            Vx, Vy = ebx * q, eby * q
            R2x, R2y = Rx - Vx, Ry - Vy
            magR2 = np.sqrt(R2x * R2x + R2y * R2y)
            s1 = (magR + magR2).astype(np.float64)
            s2 = q.astype(np.float64)
            bSq2 = (s1 * s1 - s2 * s2).astype(np.float32)
            b = f(0.5) * np.sqrt(np.maximum(bSq2, f(0)))
            act = inrange & ~(magR <= ri + rj) & (bSq2 > f(0)) & ~(b < f(0.001))
            # in_view(v_i, -R): float magnitudes and dot, double quotient
            vm = np.sqrt(V[:, 0] * V[:, 0] + V[:, 1] * V[:, 1])
            lengths = (vm[:, None] * magR).astype(np.float64)
            dt_ = (V[:, None, 0] * (-Rx) + V[:, None, 1] * (-Ry)).astype(np.float64)
            frac = dt_ / np.where(lengths == 0.0, 1.0, lengths)
            # angle() is 0 when the float cosine falls outside [-1, 1] (vector2D.h:177-190):
            # a neighbour straight behind can count as in view
            view_oob = (lengths == 0.0) | (frac < -1.0) | (frac > 1.0)
            view = view_oob | (frac > VIEW_COS)
        else:
            Rs = rj + ri
            vx = V[:, None, 0] - V[idx, 0]
            vy = V[:, None, 1] - V[idx, 1]
            a = vx * vx + vy * vy
            b = -(Rx * vx + Ry * vy)
            cc = (Rx * Rx + Ry * Ry) - Rs * Rs
            d = b * b - a * cc
            sd = np.sqrt(np.maximum(d, f(0)))
            tau = (b - sd) / a
            act = inrange & (cc > f(0)) & (d > f(0)) & ~(tau < f(0.001)) & ~(tau > f(p1))
            view = np.ones_like(act)
            view_oob = np.zeros_like(act)
        inter = (ri + rj) - magR
        contact = mask & (inter > f(0)) & (magR > f(0))
    t = torch.from_numpy
    return {"reached": t(reached), "inrange": t(inrange), "act": t(act), "view": t(view), "contact": t(contact),
            "pref32": t(np.stack([prx, pry], -1).astype(np.float64)), "view_oob": t(view_oob)}


def f64_decisions(pos, vel, idx, off, mask, c, p1):
    """The same branch decisions taken on the float64 values of the step."""
    with torch.no_grad():
        pos, vel = pos.detach(), vel.detach()
        radius = c["radius"].to(pos.device)
        d = c["goal"].to(pos.device) - pos
        reached = (d * d).sum(-1) <= radius ** 2
        dn, _ = _safe_normalize(d)
        pref = torch.where(reached[:, None], torch.zeros_like(d), dn * c["pref_speed"].to(pos.device)[:, None])
        pj = pos[idx] + off
        Rv = pos[:, None, :] - pj
        dsq = (Rv * Rv).sum(-1)
        magR = _norm(Rv)
        ri, rj = radius[:, None], radius[idx]
        inrange = mask & (dsq <= c["rng"] ** 2)
        if c["kind"] == KIND_SF:
            vb = _norm(vel[idx])
            Eb, _ = _safe_normalize(pref[idx])
            V = (vb * c["sdt"])[..., None] * Eb
            magR2 = _norm(Rv - V)
            bSq2 = (magR + magR2) ** 2 - (vb * c["sdt"]) ** 2
            act = inrange & (magR > ri + rj) & (bSq2 > 0)
            act = act & (0.5 * torch.sqrt(torch.where(act, bSq2, torch.ones_like(bSq2))) >= 0.001)
            # view factor: angle(v_i, -R) < 100 degrees
            vi = vel[:, None, :].expand_as(Rv)
            lv = _norm(vi) * magR
            cosv = torch.where(lv > 0, (vi * (-Rv)).sum(-1) / torch.where(lv > 0, lv, torch.ones_like(lv)),
                               torch.ones_like(lv))
            view = cosv > VIEW_COS
        else:
            v = vel[:, None, :] - vel[idx]
            a = (v * v).sum(-1)
            bq = -(Rv * v).sum(-1)
            cq = (Rv * Rv).sum(-1) - (ri + rj) ** 2
            dq = bq * bq - a * cq
            okd = inrange & (cq > 0) & (dq > 0)
            tau = (bq - torch.sqrt(torch.where(okd, dq, torch.ones_like(dq)))) / torch.where(okd, a, torch.ones_like(a))
            act = okd & (tau >= 0.001) & (tau <= float(p1))
            view = torch.ones_like(act)
        contact = mask & (ri + rj - magR > 0) & (magR > 0)
    return {"reached": reached, "inrange": inrange, "act": act, "view": view, "contact": contact}


# ---------------------------------------------------------------- the step
def _norm(v):
    """|v| with a finite gradient at v = 0 (where torch.where masks are applied)."""
    s = (v * v).sum(-1)
    ok = s > 0
    return torch.where(ok, torch.sqrt(torch.where(ok, s, torch.ones_like(s))), torch.zeros_like(s))


def _safe_normalize(v):
    m = _norm(v)
    ok = m > 0
    return torch.where(ok[..., None], v / torch.where(ok, m, torch.ones_like(m))[..., None], torch.zeros_like(v)), m


def _clamp(v, L, clamped=None):
    """clampVector (vector2D.h:241-247); `clamped` gives the branch instead of |v| > L."""
    m = _norm(v)
    big = m > L if clamped is None else clamped
    scale = torch.where(big, L / torch.where(big, m, torch.ones_like(m)), torch.ones_like(m))
    return v * scale[..., None]


def step(pos, vel, p0, p1, c, lists=None, branches="f64", clamps=None):
    """One step. p0, p1: the interaction model's differentiated parameters, (V0, sigma)
    of SocialForcesAvoidance (the Helbing A, B) or (k, tau0) of PowerLaw. c: constants
    (consts_from). lists: (idx, off, mask) or None (brute force from pos). branches:
    "f64" or "f32" (see the module docstring); with "f32" the preferred velocities take
    the float values the FP32 step computed (their derivative is the exact one), so the
    step is linearised where the FP32 step's reverse pass linearises it. clamps: optional
    per-agent flags (bit 0 acceleration clamped, bit 1 speed clamped) that give the clamp
    branches, as the FP32 step decided them (umans_adjoint_clamps)."""
    dt = c["dt"]
    idx, off, mask = lists if lists is not None else neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"],
                                                                     c["cap"])
    dev = pos.device
    idx, off, mask = idx.to(dev), off.to(dev), mask.to(dev)
    p1v = float(torch.as_tensor(p1).detach())
    if branches == "f32":
        D = f32_decisions(pos.detach().cpu().numpy(), vel.detach().cpu().numpy(), idx.cpu().numpy(),
                          off.cpu().numpy(), mask.cpu().numpy(), c, p1v)
        D = {k: v.to(dev) for k, v in D.items()}
    else:
        D = f64_decisions(pos, vel, idx, off, mask, c, p1v)
    d = c["goal"].to(dev) - pos
    radius = c["radius"].to(dev)
    dn, _ = _safe_normalize(d)
    reached = D["reached"]
    pref = torch.where(reached[:, None], torch.zeros_like(d), dn * c["pref_speed"].to(dev)[:, None])
    if "pref32" in D:
        pref = D["pref32"] + (pref - pref.detach())
    pj = pos[idx] + off                                   # phantom positions (N, K, 2)
    vj = vel[idx]
    rj = radius[idx]
    ri = radius[:, None]
    Rv = pos[:, None, :] - pj
    magR = _norm(Rv)
    act = D["act"]
    if c["kind"] == KIND_SF:
        A, B = p0, p1
        prefj = pref[idx]
        vb = _norm(vj)
        Eb, _ = _safe_normalize(prefj)
        V = (vb * c["sdt"])[..., None] * Eb
        magV = _norm(V)
        R2 = Rv - V
        magR2 = _norm(R2)
        bSq2 = (magR + magR2) ** 2 - (vb * c["sdt"]) ** 2
        safe = lambda x: torch.where(act, x, torch.ones_like(x))
        b = 0.5 * torch.sqrt(safe(bSq2))
        cosang = torch.where(magV > 0, (Rv * V).sum(-1) / safe(magR * torch.where(magV > 0, magV, torch.ones_like(magV))),
                             torch.zeros_like(magV))
        d_magR2 = 1.0 / (2.0 * safe(magR2)) * (2.0 * magR - 2.0 * magV * cosang)
        d_bSq2 = 2.0 * (magR + magR2) * (1.0 + d_magR2)
        d_b = (1.0 / (8.0 * b)) * d_bSq2
        d_V = -A / B * torch.exp(-b / B) * d_b
        scale = torch.where(D["view"], 1.0, 0.5).to(F64)
        F = torch.where(act[..., None], -Rv * (d_V * scale)[..., None], torch.zeros_like(Rv)).sum(1)
    else:
        k, tau0 = p0, p1
        x = Rv
        v = vel[:, None, :] - vj
        a = (v * v).sum(-1)
        bq = -(x * v).sum(-1)
        cq = (x * x).sum(-1) - (ri + rj) ** 2
        dq = bq * bq - a * cq
        safe = lambda x_: torch.where(act, x_, torch.ones_like(x_))
        sd = torch.sqrt(safe(dq))
        aa = safe(a)
        tau = (bq - sd) / aa
        tau = torch.where(act, tau, torch.ones_like(tau))
        comp1 = -k * torch.exp(-tau / tau0) * (2.0 / tau + 1.0 / tau0) / (aa * tau * tau)
        comp2 = v - (aa[..., None] * x + bq[..., None] * v) / sd[..., None]
        F = torch.where(act[..., None], comp1[..., None] * comp2, torch.zeros_like(x)).sum(1)
    F_gr = (pref - vel) / max(c["relax"], dt)
    acc = c["coeff_int"] * F / c["mass_int"] + c["coeff_gr"] * F_gr / c["mass_gr"]
    acc_cl = sp_cl = None
    if clamps is not None:
        cl = torch.as_tensor(np.asarray(clamps)).to(dev)
        acc_cl, sp_cl = (cl & 1) != 0, (cl & 2) != 0
    acc = _clamp(acc, c["max_accel"].to(dev), acc_cl)
    nxt = _clamp(vel + acc * dt, c["max_speed"].to(dev), sp_cl)
    nxt = torch.where(reached[:, None], torch.zeros_like(nxt), nxt)
    # contact forces over every listed neighbour
    inter = ri + rj - magR
    cact = D["contact"]
    un = Rv / torch.where(cact, magR, torch.ones_like(magR))[..., None]
    contact = torch.where(cact[..., None], un * (c["contact"] * inter)[..., None], torch.zeros_like(un)).sum(1)
    v_new = nxt + contact / 80.0 * dt
    p_new = pos + v_new * dt
    if c["world_type"] == 1:
        W, H = c["W"], c["H"]
        x_ = torch.where(p_new[:, 0] > 0.5 * W, p_new[:, 0] - W,
                         torch.where(p_new[:, 0] < -0.5 * W, p_new[:, 0] + W, p_new[:, 0]))
        y_ = torch.where(p_new[:, 1] > 0.5 * H, p_new[:, 1] - H,
                         torch.where(p_new[:, 1] < -0.5 * H, p_new[:, 1] + H, p_new[:, 1]))
        p_new = torch.stack([x_, y_], -1)
    return p_new, v_new


def consts_from(sc):
    """Constants of step() from a calibration Scenario (one gradient policy with one
    interaction model, SocialForcesAvoidance or PowerLaw, and GoalReachingForce)."""
    a = sc.agents
    pol = sc.policies[0]
    kinds = list(pol["cost_kind"][:pol["n_cost"]])
    c_int = [i for i, kd in enumerate(kinds) if kd in (KIND_SF, KIND_PL)][0]
    c_gr = [i for i, kd in enumerate(kinds) if kd == KIND_GR][0]
    t = lambda x: torch.tensor(np.asarray(x, np.float64), dtype=F64)
    return {"goal": t(a["goal"]), "radius": t(a["radius"]), "pref_speed": t(a["pref_speed"]),
            "goal32": np.asarray(a["goal"], np.float32), "radius32": np.asarray(a["radius"], np.float32),
            "pref_speed32": np.asarray(a["pref_speed"], np.float32),
            "max_speed": t(a["max_speed"]), "max_accel": t(a["max_accel"]), "dt": float(np.float32(sc.dt)),
            "relax": float(np.float32(pol["relaxation_time"])), "kind": int(kinds[c_int]),
            "mass_int": float(np.float32(pol["cost_params"][c_int][0])),
            "mass_gr": float(np.float32(pol["cost_params"][c_gr][0])),
            "sdt": float(np.float32(pol["cost_params"][c_int][1])), "rng": float(np.float32(pol["cost_range"][c_int])),
            "contact": float(np.float32(pol["contact_force_scale"])), "world_type": sc.world_type,
            "W": float(np.float32(sc.width)), "H": float(np.float32(sc.height)), "cap": sc.max_neighbors,
            "coeff_int": float(np.float32(pol["cost_coeff"][c_int])),
            "coeff_gr": float(np.float32(pol["cost_coeff"][c_gr]))}


def minimg(d, c):
    if c["world_type"] != 1:
        return d
    W, H = c["W"], c["H"]
    x = torch.where(d[:, 0] > 0.5 * W, d[:, 0] - W, torch.where(d[:, 0] < -0.5 * W, d[:, 0] + W, d[:, 0]))
    y = torch.where(d[:, 1] > 0.5 * H, d[:, 1] - H, torch.where(d[:, 1] < -0.5 * H, d[:, 1] + H, d[:, 1]))
    return torch.stack([x, y], -1)


def rollout_loss(pos0, v0, p0, p1, c, targets):
    pos, vel = pos0, v0
    T = targets.shape[0]
    n = pos0.shape[0]
    loss = torch.zeros((), dtype=F64)
    for t in range(T):
        pos, vel = step(pos, vel, p0, p1, c)
        dd = minimg(pos - targets[t], c)
        loss = loss + (dd * dd).sum()
    return loss / (T * n)


def replay_gradient(states, lists, targets, p0, p1, c, device="cpu", branches="f32", clamps=None):
    """Float64 reverse sweep along a recorded FP32 trajectory.

    states[t] = (pos_t, vel_t) (float32 arrays, t = 0..T) as the FP32 simulation produced
    them; lists[t] = (idx, off, mask) of its step t; targets (T, n, 2). Each step is
    linearised at the recorded state with the FP32 step's branch decisions; the loss
    gradient of each step's output is taken at the recorded positions. Returns the loss
    (from the recorded positions) and dL/dp0, dL/dp1, dL/dpos0, dL/dvel0. branches="f64"
    decides the branches on the float64 values instead (for a float64 trajectory).
    clamps[t]: the clamp branches of step t as the FP32 step decided them (its
    acceleration is ill-conditioned where two agents' ellipses nearly degenerate, so a
    float64 re-evaluation can land on the other side of a limit)."""
    T = len(targets)
    n = states[0][0].shape[0]
    scale = 2.0 / (T * n)
    tg = torch.tensor(np.asarray(targets, np.float64), dtype=F64, device=device)
    loss = 0.0

    def lgrad(t):
        # gradient of |p_t - target_{t-1}|^2 / (T n) at the recorded position p_t
        p = torch.tensor(np.asarray(states[t][0], np.float64), dtype=F64, device=device)
        dd = minimg(p - tg[t - 1], c)
        return (dd * dd).sum().item(), scale * dd

    l, pbar = lgrad(T)
    loss += l
    vbar = torch.zeros_like(pbar)
    g0 = g1 = 0.0
    for t in range(T - 1, -1, -1):
        pos = torch.tensor(np.asarray(states[t][0], np.float64), dtype=F64, device=device, requires_grad=True)
        vel = torch.tensor(np.asarray(states[t][1], np.float64), dtype=F64, device=device, requires_grad=True)
        P0 = torch.tensor(float(p0), dtype=F64, device=device, requires_grad=True)
        P1 = torch.tensor(float(p1), dtype=F64, device=device, requires_grad=True)
        pn, vn = step(pos, vel, P0, P1, c, lists=lists[t], branches=branches,
                      clamps=None if clamps is None else clamps[t])
        gp, gv, gp0, gp1 = torch.autograd.grad((pn, vn), (pos, vel, P0, P1), grad_outputs=(pbar, vbar),
                                               allow_unused=True)
        pbar, vbar = gp, gv
        g0 += float(gp0) if gp0 is not None else 0.0
        g1 += float(gp1) if gp1 is not None else 0.0
        if t >= 1:
            l, lg = lgrad(t)
            loss += l
            pbar = pbar + lg
    return loss / (T * n), g0, g1, pbar.cpu().numpy(), vbar.cpu().numpy()
