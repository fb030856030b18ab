"""Finite differences against the pathwise gradient on the calibration crowd (a study,
not a test: pytest does not collect it).

    python tests/adjoint_fd_study.py [--steps 100] [--gpu]

Float64 part (CPU): the calibration crowd (10k agents, 100 x 100 m toric world,
SocialForcesAvoidance + GoalReachingForce, gradient method, dt 0.1 s) is rolled out in
float64 with the step of tests/adjoint_ref.py (float64 branches, the reference's
neighbour rules, lists from a k-d tree query filtered by the exact predicate). The
targets come from (A, B) = (2000, 0.08); at (1500, 0.12) it prints the exact pathwise
dL/dA, dL/dB (float64 reverse sweep along the float64 trajectory) and central
differences (L(p + h) - L(p - h)) / 2h of the float64 loss for a sequence of h,
together with the number of steps at which the two perturbed rollouts take different
branches somewhere (their counts of listed, active, in-view or touching pairs differ).

--gpu adds the FP32 rollout of the library: its adjoint and central differences of
its FP32 loss for h = 1e-2 .. 1e-6 relative.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch
from scipy.spatial import cKDTree

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), HERE]

import adjoint_ref as AR  # noqa: E402
from umans_amd import scenarios as S  # noqa: E402


def fast_lists(p, W, H, rng, cap=0):
    """The reference's neighbour lists (worldToric.cpp:32-79, AgentKDTree.cpp:35-59) of
    float64 positions: queries at p and at p + (+-W, 0), (0, +-H) where the disk crosses
    that border, d2 < float(r*r) strictly, self excluded, each segment in (d2, ID) order,
    segments concatenated; returns (idx, off, mask) as AR.neighbor_lists does."""
    n = len(p)
    r2 = float(np.float32(rng) * np.float32(rng))
    tree = cKDTree(p)
    disps = [(0.0, 0.0), (W, 0.0), (-W, 0.0), (0.0, H), (0.0, -H)]
    conds = [np.ones(n, bool), p[:, 0] - rng < -0.5 * W, p[:, 0] + rng > 0.5 * W,
             p[:, 1] - rng < -0.5 * H, p[:, 1] + rng > 0.5 * H]
    A, J, SG, D = [], [], [], []
    for s, ((dx, dy), cond) in enumerate(zip(disps, conds)):
        ii = np.nonzero(cond)[0]
        if len(ii) == 0:
            continue
        q = p[ii] + np.array([dx, dy])
        res = tree.query_ball_point(q, np.sqrt(r2) * (1 + 1e-9) + 1e-9)
        lens = np.array([len(r) for r in res])
        if lens.sum() == 0:
            continue
        a = np.repeat(ii, lens)
        j = np.concatenate([np.asarray(r, np.int64) for r in res if len(r)])
        qa = np.repeat(q, lens, axis=0)
        d2 = (qa[:, 0] - p[j, 0]) ** 2 + (qa[:, 1] - p[j, 1]) ** 2
        keep = (d2 < r2) & (j != a)
        A.append(a[keep]); J.append(j[keep]); SG.append(np.full(keep.sum(), s)); D.append(d2[keep])
    a, j, sg, d2 = (np.concatenate(x) for x in (A, J, SG, D))
    order = np.lexsort((j, d2, sg, a))
    a, j, sg = a[order], j[order], sg[order]
    counts = np.bincount(a, minlength=n)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pos_in_row = np.arange(len(a)) - starts[a]
    if cap > 0:
        keep = pos_in_row < cap
        a, j, sg, pos_in_row = a[keep], j[keep], sg[keep], pos_in_row[keep]
        counts = np.minimum(counts, cap)
    K = max(int(counts.max()), 1)
    idx = np.zeros((n, K), np.int64)
    off = np.zeros((n, K, 2))
    mask = np.zeros((n, K), bool)
    idx[a, pos_in_row] = j
    mask[a, pos_in_row] = True
    dd = np.array(disps)
    off[a, pos_in_row] = -dd[sg]
    return torch.from_numpy(idx), torch.from_numpy(off), torch.from_numpy(mask)


def rollout64(sc, c, p0, p1, T, record=False):
    pos = torch.tensor(sc.agents["pos"], dtype=AR.F64)
    vel = torch.tensor(sc.agents["vel"], dtype=AR.F64)
    P0, P1 = torch.tensor(p0, dtype=AR.F64), torch.tensor(p1, dtype=AR.F64)
    traj, states, lists, decs = [], [(pos.numpy(), vel.numpy())], [], []
    with torch.no_grad():
        for _ in range(T):
            lst = fast_lists(pos.numpy(), c["W"], c["H"], c["rng"], c["cap"])
            if record:
                lists.append(lst)
            d = AR.f64_decisions(pos, vel, *lst, c, p1)
            decs.append((int(d["act"].sum()), int((d["act"] & d["view"]).sum()), int(d["contact"].sum()),
                         int(lst[2].sum())))
            pos, vel = AR.step(pos, vel, P0, P1, c, lists=lst)
            traj.append(pos.numpy().copy())
            if record:
                states.append((pos.numpy(), vel.numpy()))
    return np.array(traj), states, lists, decs


def loss_of(traj, targets, c):
    d = traj - targets
    W, H = c["W"], c["H"]
    d[..., 0] = np.where(d[..., 0] > 0.5 * W, d[..., 0] - W, np.where(d[..., 0] < -0.5 * W, d[..., 0] + W, d[..., 0]))
    d[..., 1] = np.where(d[..., 1] > 0.5 * H, d[..., 1] - H, np.where(d[..., 1] < -0.5 * H, d[..., 1] + H, d[..., 1]))
    return float((d * d).sum() / (traj.shape[0] * traj.shape[1]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--gpu", action="store_true")
    args = ap.parse_args()
    T = args.steps
    sc = S.calibration()
    c = AR.consts_from(sc)
    A0, B0 = 1500.0, float(np.float32(0.12))
    t0 = time.time()
    targets, _, _, _ = rollout64(sc, c, 2000.0, float(np.float32(0.08)), T)
    base, states, lists, dec0 = rollout64(sc, c, A0, B0, T, record=True)
    L0 = loss_of(base, targets, c)
    l_, gA, gB, _, _ = AR.replay_gradient(states, lists, targets, A0, B0, c, branches="f64")
    print(f"# float64 rollout, {sc.n} agents, {T} steps ({time.time() - t0:.0f} s): L = {L0:.10g}; "
          f"pathwise dL/dA = {gA:.6g}, dL/dB = {gB:.6g}", flush=True)
    print("| parameter | h (relative) | FD (float64) | FD / pathwise | steps whose counts of listed, active, in-view "
          "or touching pairs differ between L(p + h) and L(p - h) |")
    print("|---|---|---|---|---|")
    for name, x0, g in (("B", B0, gB), ("A", A0, gA)):
        for hr in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10):
            h = hr * x0
            args_p = (A0, B0 + h) if name == "B" else (A0 + h, B0)
            args_m = (A0, B0 - h) if name == "B" else (A0 - h, B0)
            tp, _, _, dp = rollout64(sc, c, *args_p, T)
            tm, _, _, dm = rollout64(sc, c, *args_m, T)
            fd = (loss_of(tp, targets, c) - loss_of(tm, targets, c)) / (2 * h)
            changed = sum(1 for u, v in zip(dp, dm) if u != v)
            print(f"| {name} | {hr:.0e} | {fd:.6g} | {fd / g:.4f} | {changed} of {T} steps |", flush=True)
    if args.gpu:
        from umans_amd import adjoint as ADJ, gpu
        w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
        tg = ADJ.rollout_targets(w, sc, 2000.0, 0.08, T)

        def L32(a, b):
            w.set_policies([S.with_sf_params(p, a, b) for p in sc.policies])
            w.upload(sc.agents)
            return w.adjoint_forward(tg, T, 10)
        L32(1500.0, 0.12)
        dp, _, _ = w.adjoint_backward(1)
        print(f"# FP32 library rollout: adjoint dL/dA = {dp[0, 0]:.6g}, dL/dB = {dp[0, 1]:.6g}")
        print("| parameter | h (relative) | FD (FP32 library) | FD / adjoint |")
        print("|---|---|---|---|")
        for name, x0, g in (("B", 0.12, dp[0, 1]), ("A", 1500.0, dp[0, 0])):
            for hr in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
                h = hr * x0
                lp = L32(1500.0, 0.12 + h) if name == "B" else L32(1500.0 + h, 0.12)
                lm = L32(1500.0, 0.12 - h) if name == "B" else L32(1500.0 - h, 0.12)
                fd = (lp - lm) / (2 * h)
                print(f"| {name} | {hr:.0e} | {fd:.6g} | {fd / g:.4f} |", flush=True)
        w.close()


if __name__ == "__main__":
    main()
