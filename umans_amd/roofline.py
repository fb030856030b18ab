"""Frozen roofline constants for the per-step agent loop (MI355X, one GPU).

Algorithmic bytes of the fused step kernel per agent-step: it reads the agent's state
once (position 8 B, velocity 8 B, goal 8 B, radius / preferred speed / maximum speed /
maximum acceleration 16 B, policy 4 B = 44 B) and writes position, velocity and viewing
direction (24 B): 68 B. The grid rebuild adds one write and one read of an 8 B
(cell key, agent) pair: 16 B. Neighbour gathers are not counted: in cell order they
are served from cache. Total 84 B per agent-step.

Algorithmic FP32 operations per agent-step (add, mul, div, sqrt, exp, acos count 1),
counted from the reference formulas at K neighbours in range: TTC 36, PowerLaw pair 45,
SocialForces pair 80, ORCA line 40 per pair; candidate generation 10 and force-based
cost 8 per candidate and cost; preferred velocity 10, method and clamps 25, contact
12 per neighbour, integration 20 per agent.
"""

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X vector FP32 (dense)

STEP_BYTES = 68
GRID_BYTES = 16
AGENT_STEP_BYTES = STEP_BYTES + GRID_BYTES


def flops_powerlaw_sampling(k, samples, n_costs=2):
    return k * 45 + samples * (10 + 8 * n_costs) + 10 + 25 + 12 * k + 20


def flops_socialforces_gradient(k):
    return k * 80 + 10 + 25 + 12 * k + 20


def flops_orca(k):
    return k * 40 + 500 + 10 + 25 + 12 * k + 20


def flops_mixed(k=12.6):
    """Large mixed crowd: 40% PowerLaw + sampling (12 x 6), 30% SocialForces + gradient,
    30% ORCA; about 12.6 neighbours in a 4 m range at 0.25 agents per square metre."""
    return 0.4 * flops_powerlaw_sampling(k, 72) + 0.3 * flops_socialforces_gradient(k) + 0.3 * flops_orca(k)


def flops_karamouzas_sampling(k=16, samples=129):
    return samples * (50 + 40 * k) + 10 + 25 + 12 * k + 20
