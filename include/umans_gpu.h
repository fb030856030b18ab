/* umans_gpu.h — C ABI of the MI355X (gfx950) backend for the UMANS per-step agent loop.
 *
 * One call of umans_step() performs what the reference's WorldBase::DoStep does
 * (src/Engine/core/worldBase.cpp:79-128) for every agent of a world, on one GPU:
 * neighbour search (replacing AgentKDTree + World*::ComputeNeighbors), preferred
 * velocity (agent.cpp:52-59), Policy::ComputeNewVelocity (policy.cpp:43-76) with the
 * cost functions named by their registry names (costFunctionFactory.cpp:41-61),
 * contact forces (policy.cpp:107-141) and integration with toric wrap
 * (agent.cpp:71-93, worldToric.cpp:81-108).
 *
 * Conventions
 *   - Every function returns 0 on success or a negative UMANS_E* code; the message of
 *     the last failure on a context is available from umans_last_error(). Nothing
 *     throws across this boundary (the reference's C API uses bool/nullptr + stderr,
 *     src/Library/APIFunctions.h:54-112).
 *   - Host arrays are caller-owned and copied in or out; device buffers belong to the
 *     context. 2-vectors are interleaved (x0, y0, x1, y1, ...).
 *   - Calls on one context are serialised by the caller; umans_step() is asynchronous
 *     on the context's stream; downloads synchronise.
 */
#ifndef UMANS_GPU_H
#define UMANS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UMANS_ABI_VERSION 1

enum {
	UMANS_OK = 0,
	UMANS_ENOTFOUND = -2,  /* no agent with that ID */
	UMANS_EINVAL = -22,    /* bad argument */
	UMANS_ENOMEM = -12,    /* device allocation failed */
	UMANS_EHIP = -1000,    /* HIP runtime error (message in umans_last_error) */
	UMANS_ESTATE = -1001,  /* call out of order (e.g. step before upload) */
	UMANS_ETRUNC = -1002   /* a step kept other neighbours than the reference would: an agent
	                          with no cap, or with more than 32 obstacle edges in range, had
	                          more than 4096 neighbours in range (reported, and cleared, by the
	                          next umans_synchronize, umans_get_stats or umans_download; their
	                          outputs are still written) */
};

/* Cost-function kinds. Registry names (costFunctionFactory.cpp:47-60):
 * GoalReachingForce, PowerLaw, SocialForcesAvoidance, ORCA, Karamouzas, RVO,
 * Moussaid, TtcaDca, Paris, VanToll. */
enum {
	UMANS_COST_GOAL_REACHING = 0, UMANS_COST_POWER_LAW = 1, UMANS_COST_SOCIAL_FORCES = 2,
	UMANS_COST_ORCA = 3, UMANS_COST_KARAMOUZAS = 4, UMANS_COST_RVO = 5,
	UMANS_COST_MOUSSAID = 6, UMANS_COST_TTCA_DCA = 7, UMANS_COST_PARIS = 8,
	UMANS_COST_VAN_TOLL = 9, UMANS_COST_KIND_COUNT = 10
};

/* Policy::OptimizationMethod (policy.h:100-110) */
enum { UMANS_METHOD_GRADIENT = 0, UMANS_METHOD_SAMPLING = 1, UMANS_METHOD_GLOBAL = 2 };

#define UMANS_MAX_COSTS 4
#define UMANS_MAX_PARAMS 8

/* One Policy (policy.h:89-181) in packed form. The sampling fields are
 * SamplingParameters (policy.h:36-87): s_type 0 regular / 1 random; s_base 0 zero /
 * 1 current velocity; s_base_dir 0 unit / 1 current velocity / 2 preferred velocity;
 * s_radius 0 preferred speed / 1 maximum speed / 2 maximum acceleration.
 * cost_params[c] holds the cost function's own parameters in the order given by
 * umans_cost_param_name(); cost_range[c] is CostFunction::range_ (costFunction.h:132). */
typedef struct {
	int32_t method;
	int32_t s_type, s_base, s_base_dir, s_radius;
	float s_angle;
	int32_t s_speed_samples, s_angle_samples, s_random_samples, s_include_base;
	float relaxation_time, contact_force_scale;
	int32_t n_cost;
	int32_t cost_kind[UMANS_MAX_COSTS];
	float cost_coeff[UMANS_MAX_COSTS];
	float cost_range[UMANS_MAX_COSTS];
	float cost_params[UMANS_MAX_COSTS][UMANS_MAX_PARAMS];
	/* Extension, not in the reference (gradient method only; XML attributes
	 * GradientIterations and GradientStepSize on <Policy>, which the reference's loader
	 * ignores). With gradient_iterations >= 2 the method descends
	 *   v <- v - (gradient_step * dt) * sum_c coeff_c grad C_c(v)
	 * that many times from the current velocity, each gradient evaluated at the current
	 * iterate as CostFunction::GetGradient evaluates it, and takes the acceleration
	 * (v_N - v) / dt before the clamps. 0 or 1 keeps the reference's single step,
	 * acceleration = -sum_c coeff_c grad C_c(v) (policy.cpp:78-89); one iteration with
	 * gradient_step 1 is that step up to rounding. */
	int32_t gradient_iterations;
	float gradient_step;
} umans_policy;

/* World description. world_type 0 = WorldInfinite (worldInfinite.cpp), 1 = WorldToric
 * (worldToric.cpp) centred on the origin. max_neighbors caps each agent's neighbour
 * list at the first k entries of the reference's list in canonical order (ascending
 * double d2, ties by agent ID; toric segments concatenated); 1..128, or <= 0 for no
 * cap: every neighbour in range, as AgentKDTree::FindNeighbors returns them
 * (AgentKDTree.cpp:35-59). Uncapped lists of up to 4096 entries are exact (agents with
 * more than 128 are ranked by a wave each in LDS); a longer one is truncated and the
 * step reports UMANS_ETRUNC. cell_size 0 = automatic (max interaction range);
 * block_size 0 = automatic. */
typedef struct {
	int32_t world_type;
	float width, height;
	float delta_time;
	int32_t max_neighbors;
	float cell_size;
	int32_t block_size;
	int32_t device;            /* HIP device ordinal */
} umans_world_desc;

/* Per-agent host arrays (length n). Inputs for umans_upload; outputs for
 * umans_download (any output pointer may be null). rng is the per-agent
 * std::minstd_rand0 state (agent.cpp:38); null on upload = seeded from the ID. */
typedef struct {
	int32_t* id;
	float* pos;
	float* vel;
	float* goal;
	float* viewdir;
	float* radius;
	float* pref_speed;
	float* max_speed;
	float* max_accel;
	int32_t* policy;
	uint32_t* rng;
	float* pref_vel;           /* download only: preferred velocity of the last step */
	int32_t* world;            /* world of each agent in a batch (null: world 0) */
} umans_agents;

typedef struct {
	int64_t steps;             /* steps run on this context */
	int64_t overflow_agents;   /* agents with more neighbours in range than their list keeps
	                              (the cap; with no cap 4096), last step */
	int64_t max_list;          /* longest uncapped neighbour list, last step */
	double sim_time;           /* WorldBase::time_ */
	int64_t obstacle_overflow_agents;   /* agents with more obstacle edges in range than kept, last step */
} umans_stats;

typedef struct umans_ctx umans_ctx;

/* Context lifetime. Replaces CrowdSimulator construction (crowdSimulator.cpp:30-35). */
int umans_create(const umans_world_desc* desc, umans_ctx** out);
void umans_destroy(umans_ctx* ctx);
const char* umans_last_error(const umans_ctx* ctx);

/* Batched worlds: n_worlds independent worlds (each with its own type, size and
 * delta time) stepped together, every phase one launch over all their agents. The
 * reference runs one world per CrowdSimulator (crowdSimulator.h:37); a batch is the
 * same as n_worlds simulators stepped in lock-step. Neighbour cap and block size are
 * taken from worlds[0]. */
int umans_create_batch(const umans_world_desc* worlds, int n_worlds, umans_ctx** out);
int umans_num_worlds(const umans_ctx* ctx);

/* Policies (CrowdSimulator::AddPolicy, crowdSimulator.h:91). Policy i gets ID i. */
int umans_set_policies(umans_ctx* ctx, const umans_policy* policies, int n);

/* Agents: replaces WorldBase::AddAgent/Agent::setGoal (worldBase.cpp:161-187,
 * agent.cpp:120-127) for the whole population at once. */
int umans_upload(umans_ctx* ctx, const umans_agents* agents, int n);
int umans_download(umans_ctx* ctx, umans_agents* agents);
int umans_num_agents(const umans_ctx* ctx);

/* nsteps × WorldBase::DoStep (worldBase.cpp:79-128); asynchronous. */
int umans_step(umans_ctx* ctx, int nsteps);
int umans_synchronize(umans_ctx* ctx);
int umans_get_stats(umans_ctx* ctx, umans_stats* out);

/* Capture one step (every kernel, counters cleared by kernels, no memsets) in a
 * hipGraph and replay it for each step. Off by default; not used while profiling. */
int umans_set_graph(umans_ctx* ctx, int enable);

/* Launch tuning: the cell width as a factor of the largest interaction range and
 * the step-kernel block size. Neither changes any result (neighbour lists are in
 * canonical order whatever the grid). umans_autotune times trial_steps steps per
 * candidate from a saved copy of the state, restores the state, and keeps the
 * fastest; it does not advance the simulation. */
typedef struct {
	float cell_factor;
	int32_t block_size;
	float ms_per_step;         /* autotune: time of the chosen configuration */
	int32_t list_capacity;     /* lists longer than 32 (no cap, or a cap above 32): 1 (the
	                              default): every agent's neighbours collected unsorted in
	                              device memory and ranked by its wave, agents with more than
	                              128 in range redone by a 128-entry sorted insertion; 32 or
	                              64: capacity of a first sorted-insertion launch, the agents
	                              whose list is longer being redone at the next capacity;
	                              128: a single sorted-insertion launch at full capacity;
	                              0 in umans_set_tuning: unchanged. Results do not depend on
	                              it. */
	int32_t class_block[5];    /* block size of each policy class's step kernel (force +
	                              gradient, force + sampling, ORCA, generic, cooperative
	                              sampling), 0 = block_size; the autotuner refines them one
	                              class at a time after choosing the rest. The cooperative
	                              class always runs 64. Results do not depend on them. */
} umans_tuning;
int umans_autotune(umans_ctx* ctx, int trial_steps, umans_tuning* out);
int umans_set_tuning(umans_ctx* ctx, const umans_tuning* tuning);

/* Population changes of a single-world context (UMANS_EINVAL for a batch), applied on
 * the device at the next step or read-back. Rows of umans_download are the list order
 * of WorldBase::agents_: appended on insertion, the last row moves into a removed one.
 *
 * umans_add_agents: WorldBase::AddAgent (worldBase.cpp:161-187). a->id holds desired
 *   IDs (NULL or < 0: automatic); an ID already in the list is replaced by the next
 *   unused one; the IDs given are written to out_ids (may be NULL). a->goal NULL = the
 *   position (APIFunctions.cpp:163-164); a->viewdir NULL = looking at the goal;
 *   a->vel NULL = zero. start_time (NULL = now): later agents are scheduled and join
 *   the list at the first step whose start time has come (worldBase.cpp:81-86).
 *   remove_at_goal (NULL = none): the agent leaves the list after the step in which it
 *   reaches its goal (worldBase.cpp:124-127). flags: UMANS_ADD_KEEP_IDS takes a->id as
 *   final, for agents whose IDs the scenario loader assigned at load time.
 * umans_remove_agent: WorldBase::RemoveAgent (worldBase.cpp:196-235); UMANS_ENOTFOUND.
 * umans_set_agent_goal: Agent::setGoal, which also turns the agent towards the goal
 *   (agent.cpp:120-127); UMANS_ENOTFOUND.
 * umans_set_agent_states: SetAgentPositions (APIFunctions.cpp:117-137): position,
 *   velocity and viewing direction by ID (n x 2 floats each); unknown IDs are skipped.
 * umans_set_remove_at_goal: the remove_at_goal flag of listed agents. */
enum { UMANS_ADD_KEEP_IDS = 1 };   /* IDs already assigned (umans_scenario_get): used as given */
int umans_add_agents(umans_ctx* ctx, const umans_agents* agents, int n, const float* start_time,
                     const uint8_t* remove_at_goal, int flags, int32_t* out_ids);
int umans_remove_agent(umans_ctx* ctx, int32_t id);
int umans_set_agent_goal(umans_ctx* ctx, int32_t id, float x, float y);
int umans_set_agent_states(umans_ctx* ctx, int n, const int32_t* ids, const float* pos, const float* vel,
                           const float* viewdir);
int umans_set_remove_at_goal(umans_ctx* ctx, int n, const int32_t* ids, const uint8_t* flags);
int umans_num_scheduled(const umans_ctx* ctx);   /* agents waiting for their start time */

/* The list in row order as packed records with the layout of the reference's AgentData
 * (APIFunctions.h:33-49; GetAgentPositions, APIFunctions.cpp:86-115): packed on the
 * device and copied once. `out` holds umans_num_agents() records after pending
 * population changes are applied (pinned host memory copies fastest). */
typedef struct {
	int32_t id;
	float position_x, position_y;
	float velocity_x, velocity_y;
	float viewdir_x, viewdir_y;
} umans_agent_record;
int umans_download_records(umans_ctx* ctx, umans_agent_record* out, int capacity);

/* Obstacles of a single-world context (UMANS_EINVAL for a batch): polygons given as
 * vertex counts and interleaved x,y vertices, as <Obstacle><Point .../></Obstacle>
 * (crowdSimulator.cpp:496-514); each polygon contributes the edges vertex i -> vertex
 * (i + 1) mod n (Polygon2D.cpp:27-38). Replaces any previous obstacles; n_polygons 0
 * removes them. Every agent within range of an edge sees it: time-to-collision models,
 * SocialForcesAvoidance's obstacle force and the contact forces (costFunction.cpp:229-325,
 * SocialForcesAvoidance.cpp:112-140, policy.cpp:126-137). Every edge in range is kept:
 * agents with more than 32 are redone by a wave each with the edges in device scratch,
 * so umans_stats.obstacle_overflow_agents (edges in range not kept) stays 0. */
int umans_set_obstacles(umans_ctx* ctx, int n_polygons, const int32_t* counts, const float* points);

/* CSV trajectory output: CrowdSimulator::StartCSVOutput (crowdSimulator.cpp:37-51) and
 * CSVWriter (csvwriter.cpp:37-87). After every step the positions are copied
 * asynchronously to pinned host memory and logged per agent; a flush appends rows
 * "t,x,y" (default stream formatting) to <dir>/output_<id>.csv in ID order. With
 * flush_immediately every step is flushed. A new upload flushes and stops the output.
 * Single-world contexts only (UMANS_EINVAL for a batch). The directory is created,
 * parents included. */
int umans_csv_start(umans_ctx* ctx, const char* dirname, int flush_immediately);
int umans_csv_flush(umans_ctx* ctx);
int umans_csv_stop(umans_ctx* ctx);

/* Differentiable rollout (no counterpart in the reference). Starting from the current
 * state, umans_adjoint_forward runs nsteps steps, keeping a checkpoint of the state
 * every checkpoint_every steps, and returns the loss
 *   L = 1/(nsteps * n) * sum over steps t and agents i of |p_i(t) - target_i(t)|^2
 * (the minimum-image difference in toric worlds); targets are nsteps x n x 2 floats in
 * upload order, target t being the position after step t+1. umans_adjoint_backward
 * then recomputes each segment from its checkpoint and applies the transposed Jacobian
 * of every step in reverse — the neighbour lists, filters and clamp branches held fixed
 * as the FP32 forward step decided them, the Jacobians and adjoint vectors evaluated in
 * double — and returns dL/d(parameters) as 2 floats per policy — (V0, sigma) of
 * SocialForcesAvoidance or (k, tau0) of PowerLaw, chosen by param_kind — and
 * dL/d(initial position) and dL/d(initial velocity), n x 2 floats each in upload order
 * (either may be null). targets, d_pos0 and d_vel0 may be host or device memory
 * (unified addressing; device memory must be ready when the call is made). The state
 * is left at the end of the rollout. Every policy must use the gradient method with
 * force-model costs (PowerLaw, SocialForcesAvoidance, GoalReachingForce). */
typedef struct {
	int32_t nsteps;
	int32_t checkpoint_every;
	int32_t param_kind;        /* UMANS_COST_SOCIAL_FORCES or UMANS_COST_POWER_LAW */
} umans_adjoint_desc;
int umans_adjoint_forward(umans_ctx* ctx, const umans_adjoint_desc* desc, const float* targets, double* loss);
int umans_adjoint_backward(umans_ctx* ctx, float* d_params, float* d_pos0, float* d_vel0);
/* The clamp branches the last umans_adjoint_backward held fixed: nsteps x n bytes (step
 * order, then upload order), bit 0 = the acceleration was clamped to the maximum
 * acceleration, bit 1 = the velocity to the maximum speed (policy.cpp:72-75), as the
 * FP32 step decided them. Host or device memory. */
int umans_adjoint_clamps(umans_ctx* ctx, uint8_t* flags);

/* nsteps × WorldBase::DoStep of one world on caller-owned DEVICE arrays, allocating
 * nothing. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). It gives the results of umans_create + umans_set_policies +
 * umans_upload + umans_step(nsteps) + umans_download with the same world, policies and
 * agents, through the same kernels.
 *   agents: device pointers, row order. Inputs id, goal, radius, pref_speed,
 *     max_speed, max_accel, policy; pos, vel, viewdir are read and updated in place;
 *     rng (optional) is read and updated, null = every call seeds it from the IDs
 *     (agent.cpp:38); pref_vel (optional) receives the preferred velocity of the last
 *     step; world must be null. Policy indices outside [0, n_policies) run policy 0.
 *   Limits: n <= UMANS_NOALLOC_MAX_AGENTS, 1..16 policies, 1 <= max_neighbors <= 128
 *     (lists keep the first max_neighbors entries in canonical order), no obstacles.
 *   Scratch: one fixed device array of about 5.4 MB per GPU the library runs on, held
 *     until the process ends. Every call on a GPU uses that GPU's array and leaves the
 *     last call's contents in it: calls that overlap on one GPU (other streams, other
 *     host threads) race and must be serialised by the caller.
 *   Asynchronous on `stream` (a hipStream_t; null = the null stream): the arrays are
 *   read and written by kernels on that stream. err (optional) receives a message. */
#define UMANS_NOALLOC_MAX_AGENTS 16384
int umans_step_noalloc(const umans_world_desc* world, const umans_policy* policies, int n_policies,
                       const umans_agents* agents, int n, int nsteps, void* stream, char* err, int err_len);

/* Per-phase GPU timers (hipEvents on the context's stream), accumulated over the
 * steps run while enabled: grid = cell keys + scan + reorder (+ bounding box in
 * infinite worlds), step = the fused per-agent kernel. The reference has only a
 * wall-clock timer around a whole run (crowdSimulator.cpp:109,145-153). */
typedef struct {
	int64_t steps;
	double grid_ms;
	double step_ms;
} umans_profile;
int umans_set_profiling(umans_ctx* ctx, int enable);   /* enabling resets the totals */
int umans_get_profile(umans_ctx* ctx, umans_profile* out);  /* synchronises */

/* Neighbour lists of the last step, in upload order (the reference's
 * Agent::getNeighbors after ComputeNeighbors, agent.cpp:43-50): count[n] and, per
 * agent, up to `stride` entries of neighbour ID, double squared distance to the
 * (displaced) query point, and toric segment (0 = (0,0), 1 = +W, 2 = -W, 3 = +H,
 * 4 = -H; worldToric.cpp:64-76). Requires umans_set_record_neighbors(ctx, 1)
 * before the step. */
int umans_set_record_neighbors(umans_ctx* ctx, int enable);
int umans_dump_neighbors(umans_ctx* ctx, int stride, int32_t* count, int32_t* ids,
                         double* d2, int32_t* seg);

/* Registry (CostFunctionFactory, costFunctionFactory.h:35-80 and the parameter
 * readers of each cost function's parseParameters). Host-only; no GPU needed. */
int umans_cost_id(const char* name);                           /* -1 if unknown */
const char* umans_cost_name(int kind);
int umans_cost_param_index(int kind, const char* param);       /* -1 if unknown */
const char* umans_cost_param_name(int kind, int index);        /* null past the end */
const char* umans_cost_param_read_name(int kind, int q);      /* q-th attribute in read order */
int umans_cost_defaults(int kind, float* range, float params[UMANS_MAX_PARAMS]);
int umans_policy_init(umans_policy* p, int method);            /* reference defaults */

/* Scenario files (host only; no GPU needed). Reads the reference's XML schema
 * (CrowdSimulator::FromConfigFile, crowdSimulator.cpp:516-647: configPath redirection,
 * World, delta_time, end_time, Policies / Agents / Obstacles inline or via file=) into
 * packed policies (index = order in the file; umans_scenario_get also returns each
 * policy's XML id) and agent arrays whose policy field indexes the packed policies.
 * Agent IDs are assigned as WorldBase::AddAgent assigns them. */
typedef struct {
	int32_t world_type;
	float width, height;
	float delta_time;
	float end_time;            /* -1 when absent */
	int32_t n_agents, n_policies, n_obstacles, n_warnings;
} umans_scenario_info;
typedef struct umans_scenario umans_scenario;
int umans_scenario_load(const char* path, umans_scenario** out, char* err, int err_len);
void umans_scenario_free(umans_scenario* sc);
int umans_scenario_get_info(const umans_scenario* sc, umans_scenario_info* info);
const char* umans_scenario_warning(const umans_scenario* sc, int i);
int umans_scenario_get(const umans_scenario* sc, umans_policy* policies, int32_t* policy_xml_ids,
                       umans_agents* agents, float* start_time, int32_t* remove_at_goal);
int umans_scenario_obstacles(const umans_scenario* sc, int32_t* counts, float* points);

#ifdef __cplusplus
}
#endif
#endif
