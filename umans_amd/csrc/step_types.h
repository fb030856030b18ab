// Types shared by the host side and the kernels of libumans_gpu.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace umans {

enum {
	KIND_GOAL_REACHING = 0, KIND_POWER_LAW = 1, KIND_SOCIAL_FORCES = 2, KIND_ORCA = 3,
	KIND_KARAMOUZAS = 4, KIND_RVO = 5, KIND_MOUSSAID = 6, KIND_TTCA_DCA = 7,
	KIND_PARIS = 8, KIND_VAN_TOLL = 9
};

// How a policy turns costs into a new velocity (Policy::ComputeNewVelocity,
// policy.cpp:43-76, with getBestVelocityGlobal resolved on the host).
enum {
	MODE_GRADIENT = 0,       // a = -sum coeff * gradient (policy.cpp:78-89)
	MODE_SAMPLE = 1,         // best = argmin over candidates (costFunction.cpp:66-164)
	MODE_GLOBAL_ORCA = 2,    // best = ORCA LP solution (ORCA.cpp:88-95)
	MODE_GLOBAL_FORCE = 3    // best = target velocity of the force model (ForceBasedFunction.cpp:38-41)
};

// meta.y packs the policy (low bits) and the world of a batch (high bits)
constexpr int kPolicyBits = 10;
constexpr int kPolicyMask = (1 << kPolicyBits) - 1;

constexpr unsigned kSegShift = 28;
constexpr unsigned kIndexMask = (1u << kSegShift) - 1u;
constexpr int kStatCopies = 64;
constexpr int kStatLine = 16;
constexpr int kMaxAngleSamples = 64;
constexpr int kMaxGradientIterations = 1000;   // the gradient method's descent extension

// A policy as the kernels read it. The sampling fields are the effective ones (for a
// global policy that falls back to sampling: SamplingParameters::
// ApproximateGlobalOptimization, policy.h:74-86).
struct DevPolicy {
	int mode;
	int s_type, s_base, s_base_dir, s_radius;
	int s_speed_samples, s_angle_samples, s_random_samples, s_include_base;
	int s_angle_div;         // angle == 360 ? angle_samples : angle_samples - 1
	float max_angle;         // (float)(angle / 360.0 * PI), costFunction.cpp:90
	float relaxation, contact_scale;
	float query_range;       // Policy::getInteractionRange, policy.cpp:34-41
	int n_cost;
	int orca_cost;           // index of the ORCA cost or -1
	int cls;                 // kernel class (CLS_* in kernels.h)
	double view_cos;         // SocialForces view test: (float)acos(c) < 100 deg  <=>  c > view_cos
	int grad_iters;          // gradient method: 1 = the reference's single step, >= 2 descent
	float grad_step;         // descent step in frames (umans_policy::gradient_step)
	int kind[4];
	float coeff[4];          // effective coefficients (1 for a single-cost global fallback)
	float range[4];
	float params[4][8];
	float ang_cos[kMaxAngleSamples];   // regular sampling: cosf/sinf of the float angle
	float ang_sin[kMaxAngleSamples];   // sequence start, start + dA, ... (host libm)
};

// One world of a batch; a single world is a batch of one (the reference has one world
// per CrowdSimulator, crowdSimulator.h:37).
struct WorldInfo {
	int world_type;          // 0 infinite, 1 toric
	float W, H;
	float dt;
	int cell_offset;         // first global cell of this world's grid
	int cell_cap;            // cells reserved for this world's grid
	float cell_size;         // requested cell width (m)
	int pad;
};

struct GridInfo {
	float ox, oy;            // origin of cell (0, 0)
	float invx, invy;        // cells per metre
	int ncx, ncy;
	int ncells;
	int pad;
};

// Cell-sorted (after reorder) agent state, structure of arrays.
struct State {
	float2* pos;
	float2* vel;
	float2* goal;
	float2* viewdir;
	float2* pref;            // preferred velocity of the current step
	float4* prm;             // radius, preferred speed, maximum speed, maximum acceleration
	int4* meta;              // agent ID, policy | world << kPolicyBits, upload row, RNG state
};

struct StepArgs {
	State in;                // cell-sorted state of this step (read only)
	State out;               // next state, same order; goal, prm, pref and meta null: those stay in
	                         // `in` (only the engine state in meta.w is updated there)
	const DevPolicy* pol;
	const WorldInfo* worlds;
	const GridInfo* grids;
	const GridInfo* next_grids;   // the next step's grids (count_next); infinite worlds: set up after the reorder
	const unsigned* cell_start;   // first sorted index of each cell (all worlds), + total
	const int* class_list;        // sorted index of each class-list entry
	int base;                     // first sorted index of this launch's agents
	int n;                        // agents of this launch
	int kcap;                // neighbour cap (<= KMAX of the kernel instance)
	// optional neighbour record (null when not recording)
	int* rec_count;
	int* rec_id;
	double* rec_d2;
	int* rec_seg;
	int rec_stride;
	// optional raw lists (sorted index | segment << kSegShift, stride KMAX) for the adjoint
	unsigned* raw_nb;
	int* raw_count;
	int raw_stride;
	// step counters, kStatCopies copies of kStatLine (one 128-byte line each): [0] overflow
	// agents, [1] max list length, [2] obstacle overflow agents. Atomics on one line are
	// serialised, so each wave updates the copy of its wave index (stat_copy); the host
	// sums (0, 2) and maxes (1) the copies.
	unsigned long long* stats;
	// obstacle edges of a single-world context (Polygon2D::GetEdges order), or none
	const float4* edges;
	int n_edges;
	// list-capacity tiers: a launch may take its agents from a list (sorted indices at
	// agent_list[base + t], t < *agent_count) and may hand on the agents whose neighbour
	// list outgrew its capacity (ovf_list[ovf_base + i], i counted in *ovf_count)
	const int* agent_list;
	const unsigned* agent_count;
	// optional: the launch's first list entry read from the device instead of `base` (the
	// class ranges of a context-free step are known only on the device)
	const unsigned* dev_base;
	int* ovf_list;
	unsigned* ovf_count;
	int ovf_base;
	int tier_final;               // 1: this launch's lists are final (counts go to stats)
	// with no cap (uncapped = 1) an agent with more than 128 neighbours in range, and in
	// any world an agent with more than kMaxObstacleEdges obstacle edges in range, is
	// handed to k_step_huge: huge_list[i], i counted in *huge_count (null: no such path)
	int* huge_list;
	unsigned* huge_count;
	int uncapped;
	// sticky error bits (kFlag*), kept until the host reads and clears them
	unsigned* flags;
	// optional: the next step's grid count done here (fixed grids only): the key of every
	// agent, from its new position, written into next_cell_of (indexed like out) and
	// counted in next_count; next_key_base = this class's slot x slot cells
	unsigned* next_count;
	unsigned* next_cell_of;
	int next_key_base;
	// long lists (the step kernel instance with KMAX = 0): lane t collects its in-range
	// candidates unsorted into row base + t of long_key (entry e at
	// [row * kLongCap + e], packed as segment | float(d2) | index), at most kLongCap
	// of them; the wave then sorts each row and writes its entries, in order, into
	// long_u (index | segment; entry e at [e * long_stride + row], so a wave's lanes
	// read their e-th entries together)
	unsigned long long* long_key;
	unsigned* long_u;
	int long_stride;
	// agents per wave of the long-list instance (16, 32 or 64; 0 = 64): fewer agents per
	// wave give small crowds more waves and each wave fewer rows to rank
	int long_apw;
};

// Capacity of a collected long list; an agent with more candidates in range is handed
// to a launch with the 128-entry sorted insertion, which keeps the nearest.
constexpr int kLongCap = 128;

// Obstacle edges kept per agent by a lane of the generic class; an agent with more is
// handed to k_step_huge, which keeps them all.
constexpr int kMaxObstacleEdges = 32;

// Entries of a neighbour list k_step_huge ranks in LDS: an uncapped list longer than this
// sets kFlagListTooLong (the host reports it as an error).
constexpr int kHugeMax = 4096;
enum : unsigned { kFlagListTooLong = 1u, kFlagRawListTooLong = 2u };

} // namespace umans
