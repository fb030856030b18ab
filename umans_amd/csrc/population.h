// Device side of population changes between steps: agents added by the host, removed
// (WorldBase::RemoveAgent, remove_at_goal) or edited by ID (SetAgentPositions,
// SetAgentGoal). The state stays on the device: edits address agents through their
// upload row (meta.z), removals compact the state into the other buffer, and agents
// that reached their goal are found by a kernel that hands the host only their rows.
#pragma once

#include "step_types.h"

namespace umans {

enum : int {
	EDIT_POS = 1,
	EDIT_VEL = 2,
	EDIT_VIEW = 4,
	EDIT_GOAL = 8,
	EDIT_VIEW_FROM_GOAL = 16,   // Agent::setGoal with the device position (agent.cpp:120-127)
};

struct PopEdit {
	int row;                    // upload row on the device
	int mask;                   // EDIT_* bits
	float2 pos, vel, view, goal;
};

// inv[row] = sorted index of the agent in that upload row
__global__ __launch_bounds__(256) void k_row_index(const int4* __restrict__ meta, int n, int* __restrict__ inv) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) inv[meta[k].z] = k;
}

// One thread per edited agent (the host merges the edits of one agent, in call order).
__global__ __launch_bounds__(256) void k_apply_edits(State s, const int* __restrict__ inv,
                                                     const PopEdit* __restrict__ edits, int m) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= m) return;
	const PopEdit e = edits[i];
	const int k = inv[e.row];
	if (e.mask & EDIT_POS) s.pos[k] = e.pos;
	if (e.mask & EDIT_VEL) s.vel[k] = e.vel;
	if (e.mask & EDIT_GOAL) s.goal[k] = e.goal;
	if (e.mask & EDIT_VIEW) {
		s.viewdir[k] = e.view;
	} else if (e.mask & EDIT_VIEW_FROM_GOAL) {
		// setGoal: look towards the goal unless it is the position; Vector2D::normalize
		// divides by the float magnitude when it is positive (vector2D.h:55-63)
		const float2 p = s.pos[k];
		if (e.goal.x != p.x || e.goal.y != p.y) {
			const float dx = e.goal.x - p.x, dy = e.goal.y - p.y;
			const float mag = sqrtf(dx * dx + dy * dy);
			s.viewdir[k] = mag > 0.0f ? make_float2(dx / mag, dy / mag) : make_float2(dx, dy);
		}
	}
}

// Keep the agents whose row maps to a new row (remap[row] >= 0), packed into dst. The
// packing order is that of the atomics: no result depends on the order of the agents in
// memory (the next step sorts them and neighbour lists are canonical).
__global__ __launch_bounds__(256) void k_compact(State src, State dst, int n, const int* __restrict__ remap,
                                                 unsigned* __restrict__ counter) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	int4 m = src.meta[k];
	const int nr = remap[m.z];
	if (nr < 0) return;
	const unsigned d = atomicAdd(counter, 1u);
	m.z = nr;
	dst.pos[d] = src.pos[k];
	dst.vel[d] = src.vel[k];
	dst.goal[d] = src.goal[k];
	dst.viewdir[d] = src.viewdir[k];
	dst.pref[d] = src.pref[k];
	dst.prm[d] = src.prm[k];
	dst.meta[d] = m;
}

// Agent::hasReachedGoal (agent.cpp:99-102) for agents flagged remove_at_goal: their rows.
__global__ __launch_bounds__(256) void k_goal_reached(State s, int n, const unsigned char* __restrict__ flag,
                                                      int* __restrict__ rows, unsigned* __restrict__ count) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int row = s.meta[k].z;
	if (!flag[row]) return;
	const float2 p = s.pos[k], g = s.goal[k];
	const float r = s.prm[k].x;
	const float dx = g.x - p.x, dy = g.y - p.y;
	if (dx * dx + dy * dy <= r * r) rows[atomicAdd(count, 1u)] = row;
}

// The list in row order as 28-byte records {ID, position, velocity, viewing direction}
// (the layout of the reference's AgentData, APIFunctions.h:33-49), for one host copy.
struct AgentRecord {
	int id;
	float px, py, vx, vy, dx, dy;
};
static_assert(sizeof(AgentRecord) == 28, "AgentRecord layout");

__global__ __launch_bounds__(256) void k_pack_records(State s, int n, AgentRecord* __restrict__ out) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int4 m = s.meta[k];
	const float2 p = s.pos[k], v = s.vel[k], d = s.viewdir[k];
	out[m.z] = AgentRecord{m.x, p.x, p.y, v.x, v.y, d.x, d.y};
}

} // namespace umans
