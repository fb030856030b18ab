// libumans_gpu: context management and the C ABI of include/umans_gpu.h.
// The kernels are in kernels.h; the registry in registry.cpp.
#include "umans_gpu.h"
#include "adjoint.h"
#include "population.h"
#include "step_launch.h"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <queue>
#include <string>
#include <unordered_map>
#include <vector>

using namespace umans;

namespace {

constexpr int kMaxPolicies = 1 << kPolicyBits;
constexpr int kMaxWorlds = (1 << (31 - kPolicyBits)) - 1;
constexpr int kUncappedMax = 128;
static_assert(kUncappedMax <= kAdjListMax, "the adjoint's lanes cover the longest saved list");
constexpr int kRecordUncapped = 512;   // neighbour-record entries per agent with no cap
constexpr int kHugeBlocks = 256;       // blocks (one wave each) of the very-long-list launch

template <typename T> hipError_t dalloc(T** p, size_t count) {
	return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

struct DeviceState {
	State s{};
	hipError_t alloc(int n) {
		hipError_t e;
		if ((e = dalloc(&s.pos, n)) != hipSuccess) return e;
		if ((e = dalloc(&s.vel, n)) != hipSuccess) return e;
		if ((e = dalloc(&s.goal, n)) != hipSuccess) return e;
		if ((e = dalloc(&s.viewdir, n)) != hipSuccess) return e;
		if ((e = dalloc(&s.pref, n)) != hipSuccess) return e;
		if ((e = dalloc(&s.prm, n)) != hipSuccess) return e;
		return dalloc(&s.meta, n);
	}
	void release() {
		(void)hipFree(s.pos); (void)hipFree(s.vel); (void)hipFree(s.goal); (void)hipFree(s.viewdir);
		(void)hipFree(s.pref); (void)hipFree(s.prm); (void)hipFree(s.meta);
		s = State{};
	}
};

hipError_t copy_state(const State& dst, const State& src, int n, hipStream_t st) {
	hipError_t e;
	if ((e = hipMemcpyAsync(dst.pos, src.pos, n * sizeof(float2), hipMemcpyDeviceToDevice, st))) return e;
	if ((e = hipMemcpyAsync(dst.vel, src.vel, n * sizeof(float2), hipMemcpyDeviceToDevice, st))) return e;
	if ((e = hipMemcpyAsync(dst.goal, src.goal, n * sizeof(float2), hipMemcpyDeviceToDevice, st))) return e;
	if ((e = hipMemcpyAsync(dst.viewdir, src.viewdir, n * sizeof(float2), hipMemcpyDeviceToDevice, st))) return e;
	if ((e = hipMemcpyAsync(dst.pref, src.pref, n * sizeof(float2), hipMemcpyDeviceToDevice, st))) return e;
	if ((e = hipMemcpyAsync(dst.prm, src.prm, n * sizeof(float4), hipMemcpyDeviceToDevice, st))) return e;
	return hipMemcpyAsync(dst.meta, src.meta, n * sizeof(int4), hipMemcpyDeviceToDevice, st);
}

int kmax_bucket(int kcap) {
	if (kcap <= 12) return 12;
	if (kcap <= 16) return 16;
	if (kcap <= 32) return 32;
	if (kcap <= 64) return 64;
	return 128;
}

// Block sizes instantiated per neighbour-list capacity (the LDS list is KMAX x block
// entries of 8 B).
bool block_supported(int kmax, int bs) {
	switch (kmax) {
	case 12: case 16: return bs == 64 || bs == 128 || bs == 256;
	case 32: return bs == 64 || bs == 128;
	default: return bs == 64;
	}
}
int default_block(int kmax) { return kmax <= 32 ? 128 : 64; }

// Agents per wave of the long-list step instance: 64 while that makes at least ~4 waves
// per SIMD (1024 SIMDs), else 32 or 16 — more waves for small crowds, and fewer rows for
// each wave to rank. The lists and results do not depend on it.
int long_agents_per_wave(int n) { return n >= 262144 ? 64 : (n >= 131072 ? 32 : 16); }

void launch_step(int cls, int kmax, int bs, const StepArgs& a, hipStream_t st, int coop_kind = -1) {
	static_assert(CLS_FORCE_GRAD == 0 && CLS_FORCE_SAMPLE == 1 && CLS_ORCA == 2 && CLS_GENERIC == 3, "class numbers");
	const bool short_list = kmax == 12 || kmax == 16;
	switch (cls) {
	case CLS_COOP: launch_step_coop(kmax, a, st, coop_kind); break;
	case CLS_FORCE_GRAD: (short_list ? launch_step_c0_p0 : launch_step_c0_p1)(kmax, bs, a, st); break;
	case CLS_FORCE_SAMPLE: (short_list ? launch_step_c1_p0 : launch_step_c1_p1)(kmax, bs, a, st); break;
	case CLS_ORCA: (short_list ? launch_step_c2_p0 : launch_step_c2_p1)(kmax, bs, a, st); break;
	default: (short_list ? launch_step_c3_p0 : launch_step_c3_p1)(kmax, bs, a, st); break;
	}
}

// The cooperative class works on one wave per block.
// over: the per-class block sizes (0 = bs)
int class_block(int cls, int kmax, int bs, const int* over) {
	if (cls == CLS_COOP) return 64;                 // one wave per block (k_step_coop)
	if (over && over[cls] > 0 && block_supported(kmax, over[cls])) return over[cls];
	return bs;
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

} // namespace

struct umans_ctx {
	std::vector<umans_world_desc> worlds;   // one per world of the batch
	int device = 0;
	hipStream_t stream = nullptr;
	// the step kernels of the policy classes are independent (disjoint agent ranges, read-only
	// neighbour state): each class runs on a stream of its own, forked from and joined back
	// into `stream`, so the classes share the CUs instead of running one after another
	hipStream_t cls_stream[NUM_CLS] = {};
	hipEvent_t fork_ev = nullptr;
	hipEvent_t join_ev[NUM_CLS] = {};
	std::string err;

	std::vector<DevPolicy> hpol;
	DevPolicy* dpol = nullptr;
	float max_range = 0.0f;

	int n = 0, cap = 0;
	DeviceState X, Y;            // X: state between steps, Y: cell-sorted state of a step
	DeviceState B;               // backup of X used by the autotuner
	unsigned* cell_count = nullptr;           // per (class slot, cell)
	unsigned* cell_start = nullptr;           // class-major exclusive scan of cell_count
	unsigned* cell_first = nullptr;           // per cell over all slots: the search's ranges
	unsigned* scan_sums = nullptr;          // single-pass scan: [0] ticket, [1 + t] tile t status
	unsigned* cell_of = nullptr;
	int* class_list = nullptr;                // sorted index of each class-list entry
	unsigned long long* long_key = nullptr;   // long lists: kLongCap x cap collection buffers
	unsigned* long_u = nullptr;
	// list-capacity tiers (uncapped and large caps): agents whose list outgrew a tier,
	// per tier, in their class's range; and the counts per class and tier
	int* ovf[2] = {nullptr, nullptr};
	unsigned* ovf_count = nullptr;            // [NUM_CLS][2], then the count of the very long list
	// very long lists (k_step_huge): the agents handed on, and per block of the launch its
	// obstacle edges and ORCA half-planes; used with no cap and with obstacles
	bool uncapped = false;                    // max_neighbors <= 0: every neighbour in range
	bool huge_enabled = false;
	int* huge_list = nullptr;
	float4* huge_edges = nullptr;
	Line* huge_lines = nullptr;
	int huge_edge_cap = 0, huge_alloc = 0;
	unsigned* dflags = nullptr;               // sticky error bits (kFlag*)
	int tier_first = 3;                       // 0: tiers of 32, 64, 128 entries; 1: of 64, 128; 2: a single
	                                          // 128-entry launch; 3: long lists collected and ranked per
	                                          // wave (autotuned; lists longer than 32 only)
	int cell_alloc = 0;          // allocated cell counters (>= total_cells + 1)
	// sleeps a scan tile waits for its predecessors before it sums their counts itself
	// (UMANS_SCAN_SPIN_LIMIT in the environment at creation; negative: never wait)
	int scan_spin_limit = kScanSpinLimit;
	int counts_zero_n = -1;      // the counters are zero over this key space (-1: unknown)
	// the previous step of this run_steps call counted X for this step (cell_count,
	// cell_of): the step skips k_count
	bool keys_ready = false;
	int total_cells = 0;         // sum of the worlds' cell_cap
	WorldInfo* dworlds = nullptr;
	GridInfo* dgrids = nullptr;
	GridInfo* dgrids_next = nullptr;     // infinite worlds: the next step's grids, set up after the reorder
	int* dbbox = nullptr;
	bool any_infinite = false;
	// the grids in dgrids match dworlds and the cell reservations (fixed grids only: the
	// grids of infinite worlds follow their bounding boxes every step)
	bool grids_valid = false;
	unsigned long long* dstats = nullptr;   // kStatCopies x kStatLine, see StepArgs::stats
	float4* d_edges = nullptr;                // obstacle edges (single world)
	int n_edges = 0;
	int kcap = 0;
	// policy classes: one step-kernel instance per class over a bucket of agents
	std::vector<int> h_policy;           // policy of each uploaded agent
	int cls_count[NUM_CLS] = {};
	int coop_kind = -1;                     // the one cost of every cooperative-class policy, or -1
	int cls_off[NUM_CLS] = {};              // first sorted index of each class
	int cls_slot[NUM_CLS] = {};             // slot of each present class in the sort key
	int nslots = 1;
	int* d_cls_slot = nullptr;
	float cell_factor = 1.0f;    // cell width = factor x max interaction range (autotuned)
	int block = 0;               // step-kernel block size (0 = default; autotuned)
	int cls_block[NUM_CLS] = {}; // per-class block sizes (0 = block; autotuned)

	int record = 0;
	int* rec_count = nullptr;
	int* rec_id = nullptr;
	double* rec_d2 = nullptr;
	int* rec_seg = nullptr;
	int rec_stride = 0;
	bool have_record = false;

	// hipGraphs of one step (kernels only: counters are cleared by a fill kernel). A step
	// leaves the static fields in the buffers it read them from and swaps those buffers
	// between X and Y, so consecutive steps see two buffer layouts: one graph for each,
	// told apart by the buffers they were captured with.
	int use_graph = 0;
	struct StepGraph {
		hipGraphExec_t exec = nullptr;
		hipGraph_t graph = nullptr;
		const void* key[3] = {nullptr, nullptr, nullptr};
	};
	StepGraph graphs[2];
	int graph_next = 0;

	long long steps = 0;
	double sim_time = 0.0;

	// differentiable rollout (umans_adjoint_forward / _backward)
	int adj_steps = 0, adj_ckpt = 0, adj_param_kind = -1, adj_n = 0;
	bool adj_ready = false;
	std::vector<DeviceState> ckpts;      // state before steps 0, ck, 2ck, ...
	std::vector<DeviceState> seg_states; // sorted input state of each step of a segment
	float2* targets = nullptr;           // [steps][n], upload-row order
	unsigned* seg_nb = nullptr;          // [ck][cap][KMAX]
	int* seg_cnt = nullptr;              // [ck][cap]
	double2* adj_buf = nullptr;          // 5 x n: p, v, p', v', pref
	float2* adj_out = nullptr;           // n: an adjoint vector handed out in float
	unsigned char* adj_clamps = nullptr; // [steps][n] clamp branches of the reverse pass (row order)
	bool adj_clamps_ready = false;
	double* adj_params = nullptr;        // [copies][policies][2], see AdjArgs::adj_params
	double* dloss = nullptr;
	double last_loss = 0.0;

	// CSV trajectory output (CrowdSimulator::StartCSVOutput + CSVWriter): positions are
	// copied after each step into one of two pinned slots and logged per agent
	struct CsvEntry { double t; float x, y; };
	bool csv_on = false, csv_flush_now = false;
	std::string csv_dir;
	std::vector<std::vector<CsvEntry>> csv_log;   // per log slot
	std::vector<int32_t> csv_slot_id;             // agent ID of each log slot
	std::unordered_map<int32_t, int> csv_slot_of_id;
	std::vector<int> csv_row_slot;                // log slot of each upload row
	float2* csv_pos[2] = {nullptr, nullptr};
	int4* csv_meta[2] = {nullptr, nullptr};
	hipEvent_t csv_ev[2] = {nullptr, nullptr};
	double csv_t[2] = {0.0, 0.0};
	bool csv_pending[2] = {false, false};
	int csv_slot = 0, csv_cap = 0;

	// The agent list (WorldBase::agents_, worldBase.h:117-121): upload row = list index, so
	// downloads come back in the reference's GetAgents order. Population changes between
	// steps (AddAgent, RemoveAgent, remove_at_goal, start_time, setters) are collected on
	// the host and applied on the device at the next step or read-back.
	std::vector<int32_t> row_id;                  // agent ID of each row
	std::unordered_map<int32_t, int> row_of;      // ID -> row (built on first use)
	bool row_index_built = false;
	long long next_id = 0;                        // WorldBase::nextUnusedAgentID
	std::vector<unsigned char> h_remove;          // remove_at_goal of each row
	int n_remove = 0;
	unsigned char* d_remove = nullptr;            // per device row
	bool remove_dirty = false;
	int* d_rows = nullptr;                        // rows that reached their goal / scratch
	int* d_inv = nullptr;                         // device row -> sorted index / remap
	unsigned* d_pop_count = nullptr;
	unsigned* h_pop_count = nullptr;              // pinned
	int pop_alloc = 0;                            // capacity of d_remove / d_rows / d_inv
	struct AgentRec { float2 pos, vel, goal, view; float4 prm; int4 meta; };
	bool pop_open = false;                        // edits collected since the last sync
	std::vector<int> src;                         // per row: device row, or -1 - fresh index
	std::vector<AgentRec> fresh;                  // agents added since the last sync
	std::unordered_map<int, PopEdit> edits;       // per device row, merged in call order
	// agents scheduled for insertion (WorldBase::agentsToAdd, worldBase.h:93-104)
	struct PendingRec { AgentRec a; unsigned char remove; };
	struct PendingLater {
		bool operator()(const std::pair<int, float>& a, const std::pair<int, float>& b) const {
			return a.second > b.second;
		}
	};
	std::vector<PendingRec> pending_recs;
	std::priority_queue<std::pair<int, float>, std::vector<std::pair<int, float>>, PendingLater> pending;

	// per-phase timers: one event triple per profiled step, read back lazily
	int profiling = 0;
	std::vector<hipEvent_t> ev;
	int ev_used = 0;
	long long prof_steps = 0;
	double prof_grid_ms = 0.0, prof_step_ms = 0.0;

	int fail(int code, const std::string& msg) {
		err = msg;
		return code;
	}
	int hip(hipError_t e, const char* what) {
		if (e == hipSuccess) return UMANS_OK;
		err = std::string(what) + ": " + hipGetErrorString(e);
		return UMANS_EHIP;
	}
	void drop_graph() {
		for (StepGraph& g : graphs) {
			if (g.exec) (void)hipGraphExecDestroy(g.exec);
			if (g.graph) (void)hipGraphDestroy(g.graph);
			g = StepGraph{};
		}
	}
	// The static fields (goal, radius and speeds, preferred velocity, ID / policy / row /
	// engine state) of the state a step read (Y, cell-sorted) are those of its output (X,
	// same order): the step writes only positions, velocities and viewing directions, and
	// the buffers trade places.
	void swap_static() {
		std::swap(X.s.goal, Y.s.goal);
		std::swap(X.s.prm, Y.s.prm);
		std::swap(X.s.pref, Y.s.pref);
		std::swap(X.s.meta, Y.s.meta);
	}
	void free_adjoint() {
		for (DeviceState& d : ckpts) d.release();
		for (DeviceState& d : seg_states) d.release();
		ckpts.clear();
		seg_states.clear();
		(void)hipFree(targets); (void)hipFree(seg_nb); (void)hipFree(seg_cnt); (void)hipFree(adj_buf);
		(void)hipFree(adj_out);
		(void)hipFree(adj_clamps);
		adj_out = nullptr;
		adj_clamps = nullptr;
		adj_clamps_ready = false;
		(void)hipFree(adj_params); (void)hipFree(dloss);
		targets = nullptr; seg_nb = nullptr; seg_cnt = nullptr; adj_buf = nullptr;
		adj_params = nullptr; dloss = nullptr;
		adj_ready = false;
		adj_steps = adj_ckpt = adj_n = 0;
	}
	void free_csv_buffers() {
		for (int i = 0; i < 2; ++i) {
			if (csv_pos[i]) (void)hipHostFree(csv_pos[i]);
			if (csv_meta[i]) (void)hipHostFree(csv_meta[i]);
			if (csv_ev[i]) (void)hipEventDestroy(csv_ev[i]);
			csv_pos[i] = nullptr; csv_meta[i] = nullptr; csv_ev[i] = nullptr;
			csv_pending[i] = false;
		}
		csv_cap = 0;
	}
	void free_pop_buffers() {
		(void)hipFree(d_remove); (void)hipFree(d_rows); (void)hipFree(d_inv); (void)hipFree(d_pop_count);
		if (h_pop_count) (void)hipHostFree(h_pop_count);
		d_remove = nullptr; d_rows = nullptr; d_inv = nullptr; d_pop_count = nullptr; h_pop_count = nullptr;
		pop_alloc = 0;
	}
	void free_agents() {
		drop_graph();
		free_adjoint();
		X.release(); Y.release(); B.release();
		(void)hipFree(cell_of); (void)hipFree(ovf[0]); (void)hipFree(ovf[1]); (void)hipFree(class_list);
		(void)hipFree(long_key); (void)hipFree(long_u); (void)hipFree(huge_list);
		huge_list = nullptr;
		huge_alloc = 0;
		huge_enabled = false;
		ovf[0] = ovf[1] = nullptr;
		class_list = nullptr;
		long_key = nullptr;
		long_u = nullptr;
		(void)hipFree(rec_count); (void)hipFree(rec_id); (void)hipFree(rec_d2); (void)hipFree(rec_seg);
		cell_of = nullptr;
		rec_count = rec_id = rec_seg = nullptr;
		rec_d2 = nullptr;
		n = cap = 0;
	}
	int drain_profile() {
		for (int i = 0; i + 2 < ev_used; i += 3) {
			float g = 0.0f, s = 0.0f;
			hipError_t e = hipEventSynchronize(ev[i + 2]);
			if (e == hipSuccess) e = hipEventElapsedTime(&g, ev[i], ev[i + 1]);
			if (e == hipSuccess) e = hipEventElapsedTime(&s, ev[i + 1], ev[i + 2]);
			if (e != hipSuccess) return hip(e, "profile events");
			prof_grid_ms += g;
			prof_step_ms += s;
			++prof_steps;
		}
		ev_used = 0;
		return UMANS_OK;
	}
	hipEvent_t next_event() {
		if (ev_used == (int)ev.size()) {
			hipEvent_t e = nullptr;
			if (hipEventCreate(&e) != hipSuccess) e = nullptr;   // a null event fails its record, which is checked
			ev.push_back(e);
		}
		return ev[ev_used++];
	}
	void free_events() {
		for (hipEvent_t e : ev) (void)hipEventDestroy(e);
		ev.clear();
		ev_used = 0;
	}
	float cell_size_of(const umans_world_desc& d) const {
		return d.cell_size > 0.0f ? d.cell_size : std::max(cell_factor * max_range, 0.25f);
	}
};

namespace {

// The largest double c whose angle, as vector2D.h:177-190 computes it ((float)acos(c)),
// is not below (float)(100 deg): (float)acos(x) < 100 deg  <=>  x > view_cos.
double view_cos_threshold() {
	const float half = (float)(100.0 * 3.1415926535897 / 180.0);
	double lo = -1.0, hi = 1.0;            // (float)acos(lo) >= half > (float)acos(hi)
	while (std::nextafter(lo, hi) < hi) {
		const double mid = 0.5 * (lo + hi);
		const double m = (mid <= lo || mid >= hi) ? std::nextafter(lo, hi) : mid;
		if ((float)std::acos(m) >= half) lo = m; else hi = m;
	}
	return lo;
}

// Host-side resolution of a packed policy into what the kernels read
// (Policy::ComputeNewVelocity and getBestVelocityGlobal, policy.cpp:43-100).
int make_dev_policy(const umans_policy& p, DevPolicy& d, std::string& err) {
	std::memset(&d, 0, sizeof(d));
	if (p.n_cost < 1 || p.n_cost > UMANS_MAX_COSTS) {
		err = "a policy needs 1 to 4 cost functions";
		return UMANS_EINVAL;
	}
	if (p.method < 0 || p.method > 2) {
		err = "unknown optimization method";
		return UMANS_EINVAL;
	}
	d.n_cost = p.n_cost;
	d.orca_cost = -1;
	static const double kViewCos = view_cos_threshold();
	d.view_cos = kViewCos;
	float qr = 0.0f;
	for (int c = 0; c < p.n_cost; ++c) {
		if (p.cost_kind[c] < 0 || p.cost_kind[c] >= UMANS_COST_KIND_COUNT) {
			err = "unknown cost-function kind";
			return UMANS_EINVAL;
		}
		d.kind[c] = p.cost_kind[c];
		d.coeff[c] = p.cost_coeff[c];
		d.range[c] = p.cost_range[c];
		for (int j = 0; j < UMANS_MAX_PARAMS; ++j) d.params[c][j] = p.cost_params[c][j];
		qr = std::max(qr, p.cost_range[c]);
		if (p.cost_kind[c] == UMANS_COST_ORCA) {
			if (d.orca_cost >= 0) {
				err = "at most one ORCA cost function per policy";
				return UMANS_EINVAL;
			}
			d.orca_cost = c;
		}
	}
	d.query_range = qr;
	d.relaxation = p.relaxation_time;
	d.contact_scale = p.contact_force_scale;

	umans_policy s = p;   // effective sampling parameters
	d.grad_iters = 1;
	d.grad_step = 1.0f;
	if (p.method == UMANS_METHOD_GRADIENT) {
		d.mode = MODE_GRADIENT;
		if (p.gradient_iterations >= 2) {
			if (p.gradient_iterations > kMaxGradientIterations || !std::isfinite(p.gradient_step)) {
				err = "gradient descent needs 2.." + std::to_string(kMaxGradientIterations) +
				      " iterations and a finite step";
				return UMANS_EINVAL;
			}
			d.grad_iters = p.gradient_iterations;
			d.grad_step = p.gradient_step;
		}
	} else if (p.method == UMANS_METHOD_SAMPLING) {
		d.mode = MODE_SAMPLE;
	} else {
		const int k0 = p.cost_kind[0];
		if (p.n_cost == 1 && k0 == UMANS_COST_ORCA) d.mode = MODE_GLOBAL_ORCA;
		else if (p.n_cost == 1 && (k0 == UMANS_COST_GOAL_REACHING || k0 == UMANS_COST_POWER_LAW ||
		                           k0 == UMANS_COST_SOCIAL_FORCES)) d.mode = MODE_GLOBAL_FORCE;
		else {
			// SamplingParameters::ApproximateGlobalOptimization (policy.h:74-86)
			d.mode = MODE_SAMPLE;
			s.s_type = 0; s.s_base = 0; s.s_base_dir = 0; s.s_radius = 1;
			s.s_angle = 360.0f; s.s_angle_samples = 36; s.s_speed_samples = 11; s.s_include_base = 1;
			// CostFunction::GetGlobalMinimum samples its single cost with weight 1
			// (costFunction.cpp:56-64)
			if (p.n_cost == 1) d.coeff[0] = 1.0f;
		}
	}
	d.s_type = s.s_type; d.s_base = s.s_base; d.s_base_dir = s.s_base_dir; d.s_radius = s.s_radius;
	d.s_speed_samples = s.s_speed_samples; d.s_angle_samples = s.s_angle_samples;
	d.s_random_samples = s.s_random_samples; d.s_include_base = s.s_include_base ? 1 : 0;
	if (d.mode == MODE_SAMPLE && s.s_type == 0 && (s.s_speed_samples < 1 || s.s_angle_samples < 1)) {
		err = "regular sampling needs SpeedSamples >= 1 and AngleSamples >= 1";
		return UMANS_EINVAL;
	}
	bool all_force = true;
	for (int c = 0; c < d.n_cost; ++c)
		if (d.kind[c] != KIND_GOAL_REACHING && d.kind[c] != KIND_POWER_LAW && d.kind[c] != KIND_SOCIAL_FORCES)
			all_force = false;
	if (all_force && (d.mode == MODE_GRADIENT || d.mode == MODE_GLOBAL_FORCE)) d.cls = CLS_FORCE_GRAD;
	else if (all_force && d.mode == MODE_SAMPLE) d.cls = CLS_FORCE_SAMPLE;
	else if (d.mode == MODE_GLOBAL_ORCA) d.cls = CLS_ORCA;
	else d.cls = CLS_GENERIC;
	// regular sampling over costs that only read the neighbour list per candidate:
	// candidates spread over the lanes of a wave
	if (d.cls == CLS_GENERIC && d.mode == MODE_SAMPLE && s.s_type == 0 && s.s_angle_samples <= kMaxAngleSamples) {
		bool coop = true;
		for (int c = 0; c < d.n_cost; ++c) {
			const int kd = d.kind[c];
			coop = coop && (kd == KIND_KARAMOUZAS || kd == KIND_RVO || kd == KIND_MOUSSAID || kd == KIND_TTCA_DCA ||
			                kd == KIND_PARIS || kd == KIND_VAN_TOLL);
		}
		if (coop) d.cls = CLS_COOP;
	}
	// costFunction.cpp:90 and :122-125, in the reference's float/double mix
	const float maxAngle = (float)(s.s_angle / 360.0 * 3.1415926535897);
	d.max_angle = maxAngle;
	d.s_angle_div = (s.s_angle == 360.0f) ? s.s_angle_samples : (s.s_angle_samples - 1);
	const float start = -maxAngle, end = maxAngle;
	const float dA = (end - start) / (float)d.s_angle_div;
	float ang = start;
	for (int k = 0; k < std::min(s.s_angle_samples, kMaxAngleSamples); ++k, ang += dA) {
		d.ang_cos[k] = std::cos(ang);   // float overloads, as the reference binds them
		d.ang_sin[k] = std::sin(ang);
	}
	return UMANS_OK;
}

int check_world(const umans_world_desc& d) {
	if (d.world_type != 0 && d.world_type != 1) return UMANS_EINVAL;
	if (!(d.delta_time > 0.0f)) return UMANS_EINVAL;   // crowdSimulator.cpp:575-584
	if (d.world_type == 1 && !(d.width > 0.0f && d.height > 0.0f)) return UMANS_EINVAL;  // :194-204
	return UMANS_OK;
}

// The WorldInfo record of a world holding n_agents agents (cell_offset left 0) and the
// number of cells reserved for its grid: infinite worlds 4 per agent (at least 64), the
// bounding-box grid coarsened on the device to fit; toric worlds the fixed grid of the
// world rectangle, halved along its longer side until it has at most max(4 n, 4096) cells.
long long world_info(const umans_world_desc& d, float cell_size, int n_agents, WorldInfo& r) {
	r.world_type = d.world_type;
	r.W = d.width;
	r.H = d.height;
	r.dt = d.delta_time;
	r.cell_size = cell_size;
	r.cell_offset = 0;
	r.pad = 0;
	const long long cap_cells = std::max(4LL * n_agents, 64LL);
	if (d.world_type != 1) {
		r.cell_cap = (int)cap_cells;
		return cap_cells;
	}
	// the same reduction k_grid_setup applies on the device
	double nx = std::max(1.0, std::floor((double)d.width / (double)r.cell_size));
	double ny = std::max(1.0, std::floor((double)d.height / (double)r.cell_size));
	const double lim = (double)std::max(cap_cells, 4096LL);
	while (nx * ny > lim) {
		if (nx >= ny) nx = std::max(1.0, std::floor(nx / 2.0)); else ny = std::max(1.0, std::floor(ny / 2.0));
	}
	r.cell_cap = (int)std::max(cap_cells, 4096LL);
	return (long long)(nx * ny);
}

// Per-world cell reservations and WorldInfo records.
int setup_worlds(umans_ctx* c, const std::vector<int>& per_world) {
	const int nw = (int)c->worlds.size();
	// new grids: the next step sets them up, eagerly (the captured step assumes them)
	c->grids_valid = false;
	c->keys_ready = false;
	c->drop_graph();
	std::vector<WorldInfo> wi(nw);
	long long off = 0;
	c->any_infinite = false;
	for (int w = 0; w < nw; ++w) {
		const umans_world_desc& d = c->worlds[w];
		WorldInfo& r = wi[w];
		const long long cells = world_info(d, c->cell_size_of(d), per_world[w], r);
		if (d.world_type != 1) c->any_infinite = true;
		r.cell_offset = (int)off;
		off += cells;
		if (off > (1LL << 30)) return c->fail(UMANS_EINVAL, "grid too large");
	}
	c->total_cells = (int)off;
	if ((long long)NUM_CLS * off + 1 > (1LL << 31) - 1) return c->fail(UMANS_EINVAL, "grid too large");
	if (NUM_CLS * c->total_cells + 1 > c->cell_alloc) {
		(void)hipFree(c->cell_count); (void)hipFree(c->cell_start); (void)hipFree(c->cell_first); (void)hipFree(c->scan_sums);
		c->cell_count = c->cell_start = c->cell_first = c->scan_sums = nullptr;
		c->cell_alloc = 0;
		const int cells = NUM_CLS * c->total_cells + 1;
		int rc;
		if ((rc = c->hip(dalloc(&c->cell_count, cells), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->cell_start, cells), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->cell_first, c->total_cells + 1), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->scan_sums, ceil_div(cells, kScanTile) + 1), "alloc grid"))) return rc;
		// the scan's ticket and statuses start at zero (each step's reorder clears them)
		if ((rc = c->hip(hipMemsetAsync(c->scan_sums, 0, (ceil_div(cells, kScanTile) + 1) * sizeof(unsigned), c->stream),
		                 "clear scan"))) return rc;
		c->cell_alloc = cells;
		c->counts_zero_n = -1;
	}
	return c->hip(hipMemcpyAsync(c->dworlds, wi.data(), nw * sizeof(WorldInfo), hipMemcpyHostToDevice, c->stream),
	              "world upload");
}

// Entries per agent of the neighbour record: the list capacity, or with no cap 512 (the
// count is always the whole list's; longer lists keep their nearest 512 in the record).
int rec_stride_of(const umans_ctx* c) { return c->uncapped ? kRecordUncapped : kmax_bucket(c->kcap); }

// Buffers of the very-long-list launch (k_step_huge): with no cap, or with obstacles.
// Allocated outside any graph capture (agent upload, obstacle change); a step without
// them in place fails.
int ensure_huge(umans_ctx* c) {
	c->huge_enabled = false;
	if (!c->uncapped && c->n_edges == 0) return UMANS_OK;
	int rc;
	if (c->huge_alloc < c->cap) {
		(void)hipFree(c->huge_list);
		c->huge_list = nullptr;
		c->huge_alloc = 0;
		if ((rc = c->hip(dalloc(&c->huge_list, std::max(c->cap, 1)), "alloc long lists"))) return rc;
		c->huge_alloc = c->cap;
	}
	if (!c->huge_lines &&
	    (rc = c->hip(dalloc(&c->huge_lines, (size_t)kHugeBlocks * 2 * kHugeMax), "alloc long lists"))) return rc;
	// an edge is in range of an agent at most once per periodic image (toric: 5 images)
	bool toric = false;
	for (const umans_world_desc& d : c->worlds) toric = toric || d.world_type == 1;
	const int edge_cap = std::max(1, (toric ? 5 : 1) * c->n_edges);
	if (c->huge_edge_cap < edge_cap) {
		(void)hipFree(c->huge_edges);
		c->huge_edges = nullptr;
		c->huge_edge_cap = 0;
		if ((rc = c->hip(dalloc(&c->huge_edges, (size_t)kHugeBlocks * edge_cap), "alloc long lists"))) return rc;
		c->huge_edge_cap = edge_cap;
	}
	c->huge_enabled = c->cap > 0;
	return UMANS_OK;
}

// The sticky error bits the kernels set since the last call (the stream is synchronised);
// a set bit is cleared once reported. kFlagListTooLong: an agent of the very-long-list
// launch (no cap, or more than 32 obstacle edges in range) had more than kHugeMax
// neighbours in range and its list kept kHugeMax of them, not the nearest.
// kFlagRawListTooLong: an adjoint segment met a list longer than the adjoint keeps.
int take_flags(umans_ctx* c, unsigned mask) {
	unsigned h = 0;
	int rc;
	if ((rc = c->hip(hipMemcpyAsync(&h, c->dflags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream), "flags"))) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "flags"))) return rc;
	if (!(h & mask)) return UMANS_OK;
	const unsigned keep = h & ~mask;
	if ((rc = c->hip(hipMemcpy(c->dflags, &keep, sizeof(unsigned), hipMemcpyHostToDevice), "flags"))) return rc;
	if (h & mask & kFlagListTooLong)
		return c->fail(UMANS_ETRUNC, "an agent had more than " + std::to_string(kHugeMax) +
		                                 " neighbours in range: its list was truncated (results differ from the reference's)");
	return c->fail(UMANS_ETRUNC, "an agent had more than " + std::to_string(kUncappedMax) +
	                                 " neighbours in range: the adjoint keeps " + std::to_string(kUncappedMax));
}

// Agents per policy class from the uploaded policies (host copy); the classes present
// get consecutive slots of the sort key, so class q occupies sorted indices
// [cls_off[q], cls_off[q] + cls_count[q]).
int update_classes(umans_ctx* c) {
	c->coop_kind = -2;
	for (const DevPolicy& P : c->hpol)
		if (P.cls == CLS_COOP) c->coop_kind = (P.n_cost == 1 && (c->coop_kind == -2 || c->coop_kind == P.kind[0])) ? P.kind[0] : -1;
	if (c->coop_kind < 0) c->coop_kind = -1;
	for (int q = 0; q < NUM_CLS; ++q) c->cls_count[q] = 0;
	// with obstacles every agent runs the generic class, the only one that reads edges
	for (int p : c->h_policy) c->cls_count[c->n_edges > 0 ? CLS_GENERIC : c->hpol[p].cls]++;
	int off = 0, slot = 0;
	for (int q = 0; q < NUM_CLS; ++q) {
		c->cls_off[q] = off;
		off += c->cls_count[q];
		c->cls_slot[q] = c->cls_count[q] > 0 ? slot++ : 0;
	}
	c->nslots = std::max(slot, 1);
	return c->hip(hipMemcpyAsync(c->d_cls_slot, c->cls_slot, sizeof(c->cls_slot), hipMemcpyHostToDevice, c->stream),
	              "class slots");
}

// ------------------------------------------------------------------ CSV output
int csv_take(umans_ctx* c, int slot) {
	if (!c->csv_pending[slot]) return UMANS_OK;
	int rc = c->hip(hipEventSynchronize(c->csv_ev[slot]), "csv");
	if (rc) return rc;
	const float2* p = c->csv_pos[slot];
	const int4* m = c->csv_meta[slot];
	for (int k = 0; k < c->n; ++k) {
		const int sl = c->csv_row_slot[m[k].z];
		if (sl >= 0) c->csv_log[sl].push_back({c->csv_t[slot], p[k].x, p[k].y});
	}
	c->csv_pending[slot] = false;
	return UMANS_OK;
}

// The log slot of every row of the current agent list (slots are per agent ID, so an
// agent keeps its file when rows move after removals). Agents that share an ID log one
// position per step, the last one in list order's (the std::map of
// CrowdSimulator::RunSimulationSteps, crowdSimulator.cpp:83-93); the others get -1.
void csv_rows_changed(umans_ctx* c) {
	c->csv_row_slot.resize(c->row_id.size());
	std::unordered_map<int, int> last_row;
	for (size_t r = 0; r < c->row_id.size(); ++r) {
		const int32_t id = c->row_id[r];
		auto it = c->csv_slot_of_id.find(id);
		if (it == c->csv_slot_of_id.end()) {
			it = c->csv_slot_of_id.emplace(id, (int)c->csv_slot_id.size()).first;
			c->csv_slot_id.push_back(id);
			c->csv_log.emplace_back();
		}
		c->csv_row_slot[r] = it->second;
		auto lr = last_row.find(it->second);
		if (lr != last_row.end()) c->csv_row_slot[lr->second] = -1;
		last_row[it->second] = (int)r;
	}
}

// After a step the older pending copy is in csv_slot (the slot written next), the newer
// one in csv_slot ^ 1: log them in that order.
int csv_drain(umans_ctx* c) {
	int rc = csv_take(c, c->csv_slot);
	if (!rc) rc = csv_take(c, c->csv_slot ^ 1);
	return rc;
}

// CSVWriter::Flush (csvwriter.cpp:50-72): one file output_<id>.csv per agent, in ID
// order, rows "t,x,y" in the default stream format, appended.
int csv_flush(umans_ctx* c) {
	int rc = csv_drain(c);
	if (rc) return rc;
	std::vector<int> order(c->csv_slot_id.size());
	for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
	std::sort(order.begin(), order.end(), [&](int a, int b) { return c->csv_slot_id[a] < c->csv_slot_id[b]; });
	for (int sl : order) {
		auto& log = c->csv_log[sl];
		if (log.empty()) continue;
		std::ofstream f(c->csv_dir + "output_" + std::to_string(c->csv_slot_id[sl]) + ".csv", std::ios::app);
		for (const auto& e : log) f << e.t << "," << e.x << "," << e.y << "\n";
		log.clear();
	}
	return UMANS_OK;
}

int csv_alloc(umans_ctx* c, int n) {
	c->free_csv_buffers();
	int rc;
	for (int i = 0; i < 2; ++i) {
		if ((rc = c->hip(hipHostMalloc((void**)&c->csv_pos[i], std::max(n, 1) * sizeof(float2)), "csv pinned"))) return rc;
		if ((rc = c->hip(hipHostMalloc((void**)&c->csv_meta[i], std::max(n, 1) * sizeof(int4)), "csv pinned"))) return rc;
		if ((rc = c->hip(hipEventCreate(&c->csv_ev[i]), "csv event"))) return rc;
	}
	c->csv_cap = n;
	return UMANS_OK;
}

// After a step: copy the positions (and rows) into the next pinned slot.
int csv_after_step(umans_ctx* c) {
	int rc;
	if (c->n > c->csv_cap) {                 // the population grew: wider pinned slots
		if ((rc = csv_drain(c))) return rc;
		if ((rc = csv_alloc(c, c->n + c->n / 2))) return rc;
	}
	const int slot = c->csv_slot;
	rc = csv_take(c, slot);                  // the copy two steps ago must be logged first
	if (rc) return rc;
	if ((rc = c->hip(hipMemcpyAsync(c->csv_pos[slot], c->X.s.pos, c->n * sizeof(float2), hipMemcpyDeviceToHost,
	                                c->stream), "csv copy"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(c->csv_meta[slot], c->X.s.meta, c->n * sizeof(int4), hipMemcpyDeviceToHost,
	                                c->stream), "csv copy"))) return rc;
	if ((rc = c->hip(hipEventRecord(c->csv_ev[slot], c->stream), "csv event"))) return rc;
	c->csv_t[slot] = c->sim_time;
	c->csv_pending[slot] = true;
	c->csv_slot ^= 1;
	if (c->csv_flush_now) return csv_flush(c);
	return UMANS_OK;
}

// fuse_next: this step prepares the next one (the caller guarantees nothing changes X
// before it): on fixed grids the step kernels do its k_count; with infinite worlds the
// reorder fills the bounding boxes its grids are set up from (no k_bbox pass).
// Class kernels in launch order (the first runs on the library stream, the others on
// forked ones). Measured on the 1M-agent mixed crowd: ORCA (the fewest resident waves
// per SIMD), then PowerLaw sampling, then SocialForces gives the shortest step phase
// (0.620 ms; SocialForces first 0.648, PowerLaw first 0.737).
constexpr int kLaunchOrder[NUM_CLS] = {CLS_COOP, CLS_GENERIC, CLS_ORCA, CLS_FORCE_SAMPLE, CLS_FORCE_GRAD};

int enqueue_step(umans_ctx* c, bool events, const State* ysave = nullptr, unsigned* raw_nb = nullptr,
                 int* raw_count = nullptr, bool fuse_next = false) {
	const int n = c->n;
	const State& Ys = ysave ? *ysave : c->Y.s;
	hipStream_t st = c->stream;
	const int nw = (int)c->worlds.size();
	const int scan_n = c->nslots * c->total_cells + 1;

	hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
	if (events) {
		if (c->ev_used + 3 > 3 * 256) {      // bound the pending events
			const int rc = c->drain_profile();
			if (rc) return rc;
		}
		e0 = c->next_event(); e1 = c->next_event(); e2 = c->next_event();
		if (int rc = c->hip(hipEventRecord(e0, st), "profile event")) return rc;
	}
	// prepared by the previous step of this call: the counts (fixed grids) or the
	// bounding boxes (infinite worlds)
	const bool have_keys = c->keys_ready;
	const bool fuse = fuse_next;
	const bool fuse_box = fuse_next && c->any_infinite;
	c->keys_ready = false;
	// the scan leaves the counters zero; clear them here only when the key space changed
	// (or a step did not complete, or counted for a step that did not come)
	if (!have_keys && c->counts_zero_n != scan_n)
		hipLaunchKernelGGL(k_fill, dim3(ceil_div(c->cell_alloc, 256)), dim3(256), 0, st, c->cell_count, c->cell_alloc, 0u);
	c->counts_zero_n = -1;
	if (c->any_infinite && have_keys) {
		// the grids the previous step set up and counted this step's agents in
		hipLaunchKernelGGL(k_copy_grids, dim3(ceil_div(nw, 64)), dim3(64), 0, st, c->dgrids_next, c->dgrids, nw);
	} else if (c->any_infinite) {
		hipLaunchKernelGGL(k_bbox_init, dim3(ceil_div(nw * kBoxCopies, 256)), dim3(256), 0, st, c->dbbox, nw);
		hipLaunchKernelGGL(k_bbox, dim3(ceil_div(n, 256)), dim3(256), 0, st, c->X.s.pos, c->X.s.meta, c->dworlds, n,
		                   c->dbbox);
		hipLaunchKernelGGL(k_grid_setup, dim3(ceil_div(nw, 64)), dim3(64), 0, st, c->dworlds, c->dbbox, c->dgrids, nw,
		                   nullptr, 0, nullptr, 0, fuse_box ? 1 : 0);
	} else if (!c->grids_valid) {
		// fixed grids are set up by the first step after a change; a captured step assumes them
		hipLaunchKernelGGL(k_grid_setup, dim3(ceil_div(nw, 64)), dim3(64), 0, st, c->dworlds, c->dbbox, c->dgrids, nw,
		                   nullptr, 0, nullptr, 0, 0);
		c->grids_valid = true;
	}
	if (!have_keys)
		hipLaunchKernelGGL(k_count, dim3(ceil_div(n, 256)), dim3(256), 0, st, c->X.s.pos, c->X.s.meta, n, c->dworlds,
		                   c->dgrids, c->dpol, c->d_cls_slot, c->total_cells, c->cell_count, c->cell_of);
	// single-pass scan; it also clears the step's counters. Its ticket and statuses are
	// cleared by each step's reorder, and here at the first step of a call.
	const int nb = ceil_div(scan_n, kScanTile);
	if (!have_keys)
		hipLaunchKernelGGL(k_fill, dim3(ceil_div(nb + 1, 256)), dim3(256), 0, st, c->scan_sums, nb + 1, 0u);
	constexpr int stat_words = 2 * kStatCopies * kStatLine;
	hipLaunchKernelGGL(k_scan_lookback, dim3(nb), dim3(kScanThreads), 0, st, c->cell_count, c->cell_start, c->scan_sums,
	                   c->scan_sums + 1, scan_n, nb, (unsigned*)c->dstats, stat_words, c->ovf_count, 2 * NUM_CLS + 1,
	                   c->scan_spin_limit);
	// one class slot: the per-cell starts are the class-major ones; with several the
	// reorder also writes them
	const unsigned* cell_first = c->nslots > 1 ? c->cell_first : c->cell_start;
	const int nr = std::max(c->nslots > 1 ? std::max(n, c->total_cells + 1) : n, nb + 1);
	hipLaunchKernelGGL(k_reorder, dim3(ceil_div(nr, 256)), dim3(256), 0, st, c->X.s, Ys, n, c->cell_of, c->cell_count,
	                   c->cell_start, c->nslots, c->total_cells, c->cell_first, c->class_list, c->scan_sums, nb + 1,
	                   c->dworlds, fuse_box ? c->dbbox : nullptr);
	// infinite worlds: the next step's grids from the boxes the reorder just filled; the
	// step kernels count their agents in them
	if (fuse_box)
		hipLaunchKernelGGL(k_grid_setup, dim3(ceil_div(nw, 64)), dim3(64), 0, st, c->dworlds, c->dbbox, c->dgrids_next,
		                   nw, nullptr, 0, nullptr, 0, 1);

	// without a saved copy of the sorted state (the adjoint's), the step writes only what
	// it changes and the static fields move with the buffers (swap_static)
	const bool lean = ysave == nullptr;
	StepArgs a{};
	a.in = Ys;
	a.out = c->X.s;
	if (lean) {
		a.out.goal = nullptr;
		a.out.prm = nullptr;
		a.out.pref = nullptr;
		a.out.meta = nullptr;
	}
	a.pol = c->dpol;
	a.worlds = c->dworlds;
	a.grids = c->dgrids;
	a.next_grids = c->any_infinite ? c->dgrids_next : c->dgrids;
	a.cell_start = cell_first;
	a.class_list = c->class_list;
	a.base = 0;
	a.n = n;
	a.kcap = c->kcap;
	if (c->record && c->rec_count) {
		a.rec_count = c->rec_count; a.rec_id = c->rec_id; a.rec_d2 = c->rec_d2; a.rec_seg = c->rec_seg;
		a.rec_stride = c->rec_stride;
	}
	if (raw_nb) {
		a.raw_nb = raw_nb;
		a.raw_count = raw_count;
		a.raw_stride = kmax_bucket(c->kcap);
	}
	a.stats = c->dstats;
	a.edges = c->d_edges;
	a.n_edges = c->n_edges;
	a.tier_final = 1;
	a.flags = c->dflags;
	a.uncapped = c->uncapped ? 1 : 0;
	const bool huge = c->uncapped || c->n_edges > 0;
	if (huge) {
		if (!c->huge_enabled || c->huge_alloc < n) return c->fail(UMANS_ESTATE, "long-list buffers missing");
		a.huge_list = c->huge_list;
		a.huge_count = c->ovf_count + 2 * NUM_CLS;
	}
	if (fuse) {
		a.next_count = c->cell_count;
		a.next_cell_of = c->cell_of;
	}
	const int kmax = kmax_bucket(c->kcap);
	const int bs = c->block > 0 ? c->block : default_block(kmax);
	if (e1) {
		if (int rc = c->hip(hipEventRecord(e1, st), "profile event")) return rc;
	}
	int active = 0;
	for (int q = 0; q < NUM_CLS; ++q) active += c->cls_count[q] > 0;
	if (active > 1) {
		if (int rc = c->hip(hipEventRecord(c->fork_ev, st), "fork")) return rc;
	}
	bool first = true;
	for (int qi = 0; qi < NUM_CLS; ++qi) {
		const int q = kLaunchOrder[qi];
		if (!c->cls_count[q]) continue;
		StepArgs aq = a;
		aq.base = c->cls_off[q];
		aq.n = c->cls_count[q];
		aq.next_key_base = c->cls_slot[q] * c->total_cells;
		hipStream_t sq = st;
		if (!first) {
			sq = c->cls_stream[q];
			if (int rc = c->hip(hipStreamWaitEvent(sq, c->fork_ev, 0), "fork")) return rc;
		}
		if (q == CLS_COOP || kmax <= 32 || c->tier_first == 2) {
			launch_step(q, kmax, class_block(q, kmax, bs, c->cls_block), aq, sq, c->coop_kind);
		} else if (c->tier_first == 3) {
			// Long lists: every lane collects its neighbours unsorted into its row of the
			// list buffers and the wave ranks them; the agents with more than kLongCap in
			// range are redone by a 128-entry sorted insertion, which keeps the nearest.
			// Results are those of a single launch at full capacity.
			StepArgs al = aq;
			al.long_key = c->long_key;
			al.long_u = c->long_u;
			al.long_stride = c->cap;
			al.long_apw = long_agents_per_wave(aq.n);
			al.tier_final = 0;
			al.ovf_list = c->ovf[0];
			al.ovf_count = c->ovf_count + 2 * q;
			al.ovf_base = c->cls_off[q];
			launch_step(q, 0, 64, al, sq);
			if (!c->uncapped) {           // with no cap the long lists went to k_step_huge
				StepArgs af = aq;
				af.agent_list = c->ovf[0];
				af.agent_count = c->ovf_count + 2 * q;
				af.tier_final = 1;
				launch_step(q, 128, 64, af, sq);
			}
		} else {
			// Long lists (no cap, or a cap above 32): a first launch with 32-entry lists
			// (4x the occupancy of 128) hands the agents whose list is longer to a 64-entry
			// launch, which hands on those longer still to the final capacity. Every agent
			// is computed by the first launch whose capacity holds its whole list; the
			// lists, and so the results, are those of a single launch at full capacity.
			const int tiers[3] = {32, 64, 128};
			const int nt = kmax <= 64 ? 2 : 3;
			const int t0 = c->tier_first;
			for (int ti = t0 < nt - 1 ? t0 : nt - 1; ti < nt; ++ti) {
				StepArgs at = aq;
				if (ti > t0) {
					at.agent_list = c->ovf[(ti - 1) & 1];
					at.agent_count = c->ovf_count + 2 * q + ((ti - 1) & 1);
				}
				at.tier_final = ti == nt - 1 ? 1 : 0;
				if (!at.tier_final) {
					at.ovf_list = c->ovf[ti & 1];
					at.ovf_count = c->ovf_count + 2 * q + (ti & 1);
					at.ovf_base = c->cls_off[q];
				}
				const int tk = tiers[ti];
				launch_step(q, tk, class_block(q, tk, tk <= 32 ? bs : 64, nullptr), at, sq);
			}
		}
		if (!first) {
			if (int rc = c->hip(hipEventRecord(c->join_ev[q], sq), "join")) return rc;
		}
		first = false;
	}
	if (active > 1) {
		first = true;
		for (int qi = 0; qi < NUM_CLS; ++qi) {
			const int q = kLaunchOrder[qi];
			if (!c->cls_count[q]) continue;
			if (!first) {
				if (int rc = c->hip(hipStreamWaitEvent(st, c->join_ev[q], 0), "join")) return rc;
			}
			first = false;
		}
	}
	// the agents every class kernel handed on: their count is final once all have joined
	if (huge) {
		HugeArgs H{a, c->d_cls_slot, c->total_cells, c->huge_edges, c->huge_edge_cap, c->huge_lines};
		hipLaunchKernelGGL(k_step_huge, dim3(kHugeBlocks), dim3(64), 0, st, H);
	}
	if (e2) {
		if (int rc = c->hip(hipEventRecord(e2, st), "profile event")) return rc;
	}
	const int rc = c->hip(hipGetLastError(), "step launch");
	if (!rc) {
		c->counts_zero_n = fuse ? -1 : scan_n;
		c->keys_ready = fuse_next;
		if (lean) c->swap_static();
	}
	return rc;
}

// The buffers a captured step reads and writes.
void graph_key(const umans_ctx* c, const void** key) {
	key[0] = c->X.s.pos;
	key[1] = c->X.s.goal;
	key[2] = c->Y.s.pos;
}

// The graph of one step for the current buffers, captured if there is none yet; the
// buffers are left as before (the capture's swap_static is undone: nothing ran).
int step_graph(umans_ctx* c, bool fuse_next, hipGraphExec_t* out) {
	const void* key[3];
	graph_key(c, key);
	for (umans_ctx::StepGraph& sg : c->graphs)
		if (sg.exec && sg.key[0] == key[0] && sg.key[1] == key[1] && sg.key[2] == key[2]) {
			*out = sg.exec;
			return UMANS_OK;
		}
	umans_ctx::StepGraph& sg = c->graphs[c->graph_next];
	c->graph_next ^= 1;
	if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
	if (sg.graph) (void)hipGraphDestroy(sg.graph);
	sg = umans_ctx::StepGraph{};
	int rc = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
	if (rc) return rc;
	const State x0 = c->X.s, y0 = c->Y.s;
	const int erc = enqueue_step(c, false, nullptr, nullptr, nullptr, fuse_next);
	c->X.s = x0;
	c->Y.s = y0;
	hipGraph_t g = nullptr;
	const hipError_t e = hipStreamEndCapture(c->stream, &g);
	if (erc) {
		if (g) (void)hipGraphDestroy(g);
		return erc;
	}
	if ((rc = c->hip(e, "graph capture"))) return rc;
	sg.graph = g;
	if ((rc = c->hip(hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0), "graph instantiate"))) return rc;
	for (int i = 0; i < 3; ++i) sg.key[i] = key[i];
	*out = sg.exec;
	return UMANS_OK;
}

// ------------------------------------------------------------------ population
// The agent list of a single-world context changes between steps through WorldBase's
// AddAgent / RemoveAgent / remove_at_goal / scheduled insertion (worldBase.cpp:79-235)
// and the agent setters (agent.cpp:108-127). The host keeps the list (row -> ID) and
// collects the changes; sync_population applies them on the device in one pass.

void build_row_index(umans_ctx* c) {
	if (c->row_index_built) return;
	c->row_of.clear();
	c->row_of.reserve(c->row_id.size() * 2);
	for (size_t r = 0; r < c->row_id.size(); ++r) c->row_of[c->row_id[r]] = (int)r;
	c->row_index_built = true;
}

void open_edits(umans_ctx* c) {
	if (c->pop_open) return;
	c->src.resize(c->row_id.size());
	for (size_t r = 0; r < c->src.size(); ++r) c->src[r] = (int)r;
	c->pop_open = true;
}

int ensure_pop_buffers(umans_ctx* c, int n) {
	if (c->pop_alloc >= n && c->d_pop_count) return UMANS_OK;
	const int cap = std::max(n, 64);
	(void)hipFree(c->d_remove); (void)hipFree(c->d_rows); (void)hipFree(c->d_inv);
	c->d_remove = nullptr; c->d_rows = nullptr; c->d_inv = nullptr;
	c->pop_alloc = 0;
	int rc;
	if ((rc = c->hip(dalloc(&c->d_remove, cap), "alloc population"))) return rc;
	if ((rc = c->hip(dalloc(&c->d_rows, cap), "alloc population"))) return rc;
	if ((rc = c->hip(dalloc(&c->d_inv, cap), "alloc population"))) return rc;
	if (!c->d_pop_count && (rc = c->hip(dalloc(&c->d_pop_count, 1), "alloc population"))) return rc;
	if (!c->h_pop_count && (rc = c->hip(hipHostMalloc((void**)&c->h_pop_count, sizeof(unsigned)), "pinned"))) return rc;
	c->pop_alloc = cap;
	c->remove_dirty = true;
	return UMANS_OK;
}

// Grow the agent buffers to hold n agents, keeping the first `keep` agents of X.
int grow_agents(umans_ctx* c, int n, int keep) {
	const int cap = std::max(n, c->cap + c->cap / 2);
	DeviceState nx;
	int rc;
	if ((rc = c->hip(nx.alloc(cap), "alloc state"))) return rc;
	if (keep > 0 && (rc = c->hip(copy_state(nx.s, c->X.s, keep, c->stream), "grow"))) { nx.release(); return rc; }
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "grow"))) { nx.release(); return rc; }
	const bool rec = c->rec_count != nullptr;
	c->free_agents();                         // also drops the graph, the adjoint and the record
	c->X = nx;
	if ((rc = c->hip(c->Y.alloc(cap), "alloc state"))) return rc;
	if ((rc = c->hip(dalloc(&c->cell_of, cap), "alloc grid"))) return rc;
	if ((rc = c->hip(dalloc(&c->class_list, cap), "alloc grid"))) return rc;
	if (kmax_bucket(c->kcap) > 32) {
		if ((rc = c->hip(dalloc(&c->long_key, (size_t)kLongCap * cap), "alloc lists"))) return rc;
		if ((rc = c->hip(dalloc(&c->long_u, (size_t)kLongCap * cap), "alloc lists"))) return rc;
	}
	if ((rc = c->hip(dalloc(&c->ovf[0], cap), "alloc grid"))) return rc;
	if ((rc = c->hip(dalloc(&c->ovf[1], cap), "alloc grid"))) return rc;
	c->cap = cap;
	if (rec) {
		const int stride = rec_stride_of(c);
		if ((rc = c->hip(dalloc(&c->rec_count, cap), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_id, (size_t)cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_d2, (size_t)cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_seg, (size_t)cap * stride), "alloc record"))) return rc;
		c->rec_stride = stride;
	}
	return ensure_huge(c);
}

int upload_remove_flags(umans_ctx* c) {
	c->n_remove = 0;
	for (unsigned char f : c->h_remove) c->n_remove += f != 0;
	c->remove_dirty = false;
	if (c->n_remove == 0 || c->h_remove.empty()) return UMANS_OK;
	int rc = ensure_pop_buffers(c, (int)c->h_remove.size());
	if (rc) return rc;
	c->remove_dirty = false;
	if ((rc = c->hip(hipMemcpyAsync(c->d_remove, c->h_remove.data(), c->h_remove.size(), hipMemcpyHostToDevice,
	                                c->stream), "remove flags"))) return rc;
	return c->hip(hipStreamSynchronize(c->stream), "remove flags");
}

// Setters collected for agents on the device (keyed by device row), applied in place.
int apply_device_edits(umans_ctx* c) {
	if (c->edits.empty()) return UMANS_OK;
	const int n_dev = c->n;
	int rc = ensure_pop_buffers(c, n_dev);
	if (rc) return rc;
	std::vector<PopEdit> ed;
	ed.reserve(c->edits.size());
	for (const auto& kv : c->edits) ed.push_back(kv.second);
	c->edits.clear();
	PopEdit* d_ed = nullptr;
	if ((rc = c->hip(dalloc(&d_ed, ed.size()), "alloc edits"))) return rc;
	hipLaunchKernelGGL(k_row_index, dim3(ceil_div(n_dev, 256)), dim3(256), 0, c->stream, c->X.s.meta, n_dev, c->d_inv);
	rc = c->hip(hipMemcpyAsync(d_ed, ed.data(), ed.size() * sizeof(PopEdit), hipMemcpyHostToDevice, c->stream), "edits");
	if (!rc) {
		hipLaunchKernelGGL(k_apply_edits, dim3(ceil_div((int)ed.size(), 256)), dim3(256), 0, c->stream, c->X.s,
		                   c->d_inv, d_ed, (int)ed.size());
		rc = c->hip(hipStreamSynchronize(c->stream), "edits");
	}
	(void)hipFree(d_ed);
	return rc;
}

// Apply the collected population changes: edits of device agents, removals (compaction),
// added agents (appended), then the per-row tables.
int sync_population(umans_ctx* c) {
	if (!c->pop_open) return c->remove_dirty ? upload_remove_flags(c) : UMANS_OK;
	c->keys_ready = false;     // rows move, agents come and go: the next step counts again
	int rc;
	if (c->csv_on && (rc = csv_drain(c))) return rc;   // logged copies refer to the old rows
	const int n_dev = c->n;
	const int n_new = (int)c->row_id.size();
	bool moved = false;
	int n_keep = 0;
	for (int r = 0; r < n_new; ++r) {
		if (c->src[r] >= 0) { ++n_keep; moved |= c->src[r] != r; }
	}
	moved |= n_keep != n_dev;
	if ((!c->edits.empty() || moved) && (rc = ensure_pop_buffers(c, std::max(n_dev, n_new)))) return rc;
	// 1. setters on device agents, addressed by their device row
	if ((rc = apply_device_edits(c))) return rc;
	// 2. removed and moved device agents: compact into Y with their new rows, swap
	if (moved && n_dev > 0) {
		std::vector<int> remap(n_dev, -1);
		for (int r = 0; r < n_new; ++r)
			if (c->src[r] >= 0) remap[c->src[r]] = r;
		if ((rc = c->hip(hipMemcpyAsync(c->d_inv, remap.data(), n_dev * sizeof(int), hipMemcpyHostToDevice, c->stream),
		                 "remap"))) return rc;
		hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, c->stream, c->d_pop_count, 1, 0u);
		hipLaunchKernelGGL(k_compact, dim3(ceil_div(n_dev, 256)), dim3(256), 0, c->stream, c->X.s, c->Y.s, n_dev,
		                   c->d_inv, c->d_pop_count);
		if ((rc = c->hip(hipStreamSynchronize(c->stream), "compact"))) return rc;
		std::swap(c->X, c->Y);
	}
	// 3. added agents go after the kept ones
	if (n_new > c->cap && (rc = grow_agents(c, n_new, n_keep))) return rc;
	const int n_add = n_new - n_keep;
	if (n_add > 0) {
		std::vector<float2> pos(n_add), vel(n_add), goal(n_add), view(n_add);
		std::vector<float4> prm(n_add);
		std::vector<int4> meta(n_add);
		int j = 0;
		for (int r = 0; r < n_new; ++r) {
			if (c->src[r] >= 0) continue;
			const umans_ctx::AgentRec& a = c->fresh[-1 - c->src[r]];
			pos[j] = a.pos; vel[j] = a.vel; goal[j] = a.goal; view[j] = a.view; prm[j] = a.prm;
			meta[j] = a.meta;
			meta[j].z = r;
			++j;
		}
		const State& s = c->X.s;
		const int o = n_keep;
		if ((rc = c->hip(hipMemcpyAsync(s.pos + o, pos.data(), n_add * sizeof(float2), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipMemcpyAsync(s.vel + o, vel.data(), n_add * sizeof(float2), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipMemcpyAsync(s.goal + o, goal.data(), n_add * sizeof(float2), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipMemcpyAsync(s.viewdir + o, view.data(), n_add * sizeof(float2), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipMemcpyAsync(s.prm + o, prm.data(), n_add * sizeof(float4), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipMemcpyAsync(s.meta + o, meta.data(), n_add * sizeof(int4), hipMemcpyHostToDevice, c->stream), "add"))) return rc;
		if ((rc = c->hip(hipStreamSynchronize(c->stream), "add"))) return rc;
	}
	c->pop_open = false;
	c->src.clear();
	c->fresh.clear();
	c->edits.clear();
	// 4. the per-row tables of the new list
	if (moved || n_add > 0) {
		c->n = n_new;
		c->drop_graph();
		c->free_adjoint();
		c->have_record = false;
		if ((rc = setup_worlds(c, std::vector<int>(1, n_new)))) return rc;
		if ((rc = update_classes(c))) return rc;
		c->remove_dirty = true;
		if (c->csv_on) csv_rows_changed(c);
	}
	return c->remove_dirty ? upload_remove_flags(c) : UMANS_OK;
}

// Append an agent to the list (WorldBase::addAgentToList, worldBase.cpp:189-194).
void list_append(umans_ctx* c, const umans_ctx::AgentRec& a, unsigned char remove) {
	open_edits(c);
	build_row_index(c);
	const int r = (int)c->row_id.size();
	c->row_id.push_back(a.meta.x);
	c->row_of[a.meta.x] = r;
	c->src.push_back(-1 - (int)c->fresh.size());
	c->fresh.push_back(a);
	c->h_policy.push_back(a.meta.y & kPolicyMask);
	c->h_remove.push_back(remove);
	if (remove) c->remove_dirty = true;
}

// WorldBase::removeAgentAtListIndex (worldBase.cpp:207-235): the last agent takes the row.
void list_remove_row(umans_ctx* c, int r) {
	open_edits(c);
	build_row_index(c);
	const int last = (int)c->row_id.size() - 1;
	c->row_of.erase(c->row_id[r]);
	if (r != last) {
		c->row_id[r] = c->row_id[last];
		c->src[r] = c->src[last];
		c->h_policy[r] = c->h_policy[last];
		c->h_remove[r] = c->h_remove[last];
		c->row_of[c->row_id[r]] = r;
	}
	c->row_id.pop_back();
	c->src.pop_back();
	c->h_policy.pop_back();
	c->h_remove.pop_back();
	c->remove_dirty = true;
}

// WorldBase::DoStep, before the frame (worldBase.cpp:81-86): insert the scheduled agents
// whose start time has come.
void insert_due(umans_ctx* c) {
	while (!c->pending.empty() && (double)c->pending.top().second <= c->sim_time) {
		const int i = c->pending.top().first;
		c->pending.pop();
		list_append(c, c->pending_recs[i].a, c->pending_recs[i].remove);
	}
	if (c->pending.empty()) c->pending_recs.clear();
}

// WorldBase::DoStep, after the frame (worldBase.cpp:124-127): agents flagged
// remove_at_goal that reached their goal leave the list, highest row first.
int remove_reached(umans_ctx* c) {
	if (c->n_remove == 0 || c->n == 0) return UMANS_OK;
	hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, c->stream, c->d_pop_count, 1, 0u);
	hipLaunchKernelGGL(k_goal_reached, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->X.s, c->n, c->d_remove,
	                   c->d_rows, c->d_pop_count);
	int rc = c->hip(hipMemcpyAsync(c->h_pop_count, c->d_pop_count, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream),
	                "goal check");
	if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "goal check");
	if (rc) return rc;
	const int m = (int)*c->h_pop_count;
	if (m == 0) return UMANS_OK;
	std::vector<int> rows(m);
	if ((rc = c->hip(hipMemcpy(rows.data(), c->d_rows, m * sizeof(int), hipMemcpyDeviceToHost), "goal check"))) return rc;
	std::sort(rows.begin(), rows.end(), [](int a, int b) { return a > b; });
	for (int r : rows) list_remove_row(c, r);
	return sync_population(c);
}

int run_steps_inner(umans_ctx* c, int nsteps);

// Within one call each step also counts the agents for the next one (fixed grids) or
// fills the bounding boxes for its grids (infinite worlds), so from the second step on
// k_count (k_bbox) is skipped; a population change in between makes the next step count
// again. A captured graph is the steady-state step: it is replayed only on steps whose
// keys (boxes) are ready.
int run_steps(umans_ctx* c, int nsteps) {
	const int rc0 = run_steps_inner(c, nsteps);
	c->keys_ready = false;     // keys never carry over to other entry points
	return rc0;
}

int run_steps_inner(umans_ctx* c, int nsteps) {
	const bool single = c->worlds.size() == 1;
	c->keys_ready = false;
	for (int s = 0; s < nsteps; ++s) {
		if (single) {
			insert_due(c);
			const int rc = sync_population(c);
			if (rc) return rc;
		}
		if (c->n > 0) {
			int rc;
			if (c->use_graph && !c->profiling && c->keys_ready) {
				hipGraphExec_t ge = nullptr;
				if ((rc = step_graph(c, true, &ge))) return rc;
				if ((rc = c->hip(hipGraphLaunch(ge, c->stream), "graph launch"))) return rc;
				c->swap_static();
				c->keys_ready = true;
				c->counts_zero_n = -1;
			} else if ((rc = enqueue_step(c, c->profiling != 0, nullptr, nullptr, nullptr, true))) {
				return rc;
			}
		}
		c->steps += 1;
		c->sim_time += (double)c->worlds[0].delta_time;   // WorldBase::time_ (worldBase.cpp:122)
		if (single && c->n_remove > 0) {
			const int rc = remove_reached(c);
			if (rc) return rc;
		}
		if (c->csv_on && c->n > 0) {
			const int rc = csv_after_step(c);
			if (rc) return rc;
		}
	}
	return UMANS_OK;
}

int agents_per_world(umans_ctx* c, std::vector<int>& per_world) {
	per_world.assign(c->worlds.size(), 0);
	std::vector<int4> meta(c->n);
	int rc = c->hip(hipMemcpyAsync(meta.data(), c->X.s.meta, c->n * sizeof(int4), hipMemcpyDeviceToHost, c->stream), "meta");
	if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "meta");
	if (rc) return rc;
	for (const int4& m : meta) per_world[m.y >> kPolicyBits]++;
	return UMANS_OK;
}

// ------------------------------------------------------------------ adjoint
bool adjoint_policy_ok(const DevPolicy& p) {
	if (p.mode != MODE_GRADIENT || p.grad_iters != 1) return false;
	for (int c = 0; c < p.n_cost; ++c)
		if (p.kind[c] != KIND_GOAL_REACHING && p.kind[c] != KIND_POWER_LAW && p.kind[c] != KIND_SOCIAL_FORCES) return false;
	return true;
}

void launch_loss(umans_ctx* c, const float2* target, double* loss, double2* grad, double scale) {
	hipLaunchKernelGGL(k_loss, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->X.s, c->n, c->dworlds, target,
	                   loss, grad, scale);
}

// Doubles between copies of the adjoint's parameter accumulators: whole 128-byte lines.
int adj_param_stride(int npol) { return (2 * npol + 15) / 16 * 16; }

void fill_f(umans_ctx* c, void* p, long long nfloats) {
	hipLaunchKernelGGL(k_fill_f, dim3(ceil_div(nfloats, 256)), dim3(256), 0, c->stream, (float*)p, nfloats, 0.0f);
}

int create_impl(const umans_world_desc* descs, int nworlds, umans_ctx** out) {
	if (!descs || !out || nworlds < 1 || nworlds > kMaxWorlds) return UMANS_EINVAL;
	*out = nullptr;
	for (int w = 0; w < nworlds; ++w)
		if (check_world(descs[w])) return UMANS_EINVAL;
	const int kcap = descs[0].max_neighbors > 0 ? descs[0].max_neighbors : kUncappedMax;
	const bool uncapped = descs[0].max_neighbors <= 0;
	if (kcap > kUncappedMax) return UMANS_EINVAL;
	if (descs[0].block_size > 0 && !block_supported(kmax_bucket(kcap), descs[0].block_size)) return UMANS_EINVAL;
	umans_ctx* c = new umans_ctx();
	if (const char* sl = std::getenv("UMANS_SCAN_SPIN_LIMIT")) c->scan_spin_limit = std::atoi(sl);
	c->worlds.assign(descs, descs + nworlds);
	c->device = descs[0].device;
	c->kcap = kcap;
	c->uncapped = uncapped;
	c->block = std::max(descs[0].block_size, 0);
	bool streams_ok = hipSetDevice(c->device) == hipSuccess &&
	                  hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) == hipSuccess;
	for (int q = 0; q < NUM_CLS && streams_ok; ++q)
		streams_ok = hipStreamCreateWithFlags(&c->cls_stream[q], hipStreamNonBlocking) == hipSuccess &&
		             hipEventCreateWithFlags(&c->join_ev[q], hipEventDisableTiming) == hipSuccess;
	if (!streams_ok || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
	    dalloc(&c->dgrids, nworlds) != hipSuccess || dalloc(&c->dgrids_next, nworlds) != hipSuccess || dalloc(&c->dworlds, nworlds) != hipSuccess ||
	    dalloc(&c->dbbox, (size_t)kBoxCopies * kBoxLine * nworlds) != hipSuccess ||
	    dalloc(&c->dstats, (size_t)kStatCopies * kStatLine) != hipSuccess ||
	    dalloc(&c->ovf_count, 2 * NUM_CLS + 1) != hipSuccess || dalloc(&c->dpol, kMaxPolicies) != hipSuccess ||
	    dalloc(&c->dflags, 32) != hipSuccess ||
	    hipMemsetAsync(c->dflags, 0, 32 * sizeof(unsigned), c->stream) != hipSuccess ||
	    dalloc(&c->d_cls_slot, NUM_CLS) != hipSuccess ||
	    // the counters read before the first step: zero
	    hipMemsetAsync(c->dstats, 0, (size_t)kStatCopies * kStatLine * sizeof(unsigned long long), c->stream) != hipSuccess) {
		umans_destroy(c);
		return UMANS_EHIP;
	}
	*out = c;
	return UMANS_OK;
}

} // namespace

extern "C" {

int umans_create(const umans_world_desc* desc, umans_ctx** out) { return create_impl(desc, 1, out); }

int umans_create_batch(const umans_world_desc* worlds, int n_worlds, umans_ctx** out) {
	return create_impl(worlds, n_worlds, out);
}

int umans_num_worlds(const umans_ctx* c) { return c ? (int)c->worlds.size() : 0; }

void umans_destroy(umans_ctx* c) {
	if (!c) return;
	(void)hipSetDevice(c->device);
	if (c->stream) (void)hipStreamSynchronize(c->stream);
	if (c->csv_on) csv_flush(c);
	c->free_csv_buffers();
	c->free_pop_buffers();
	c->free_agents();
	(void)hipFree(c->dgrids); (void)hipFree(c->dgrids_next); (void)hipFree(c->dworlds); (void)hipFree(c->dbbox); (void)hipFree(c->dstats); (void)hipFree(c->dpol);
	(void)hipFree(c->d_edges); (void)hipFree(c->ovf_count); (void)hipFree(c->dflags);
	(void)hipFree(c->huge_edges); (void)hipFree(c->huge_lines);
	(void)hipFree(c->d_cls_slot);
	(void)hipFree(c->cell_count); (void)hipFree(c->cell_start); (void)hipFree(c->cell_first); (void)hipFree(c->scan_sums);
	c->free_events();
	for (int q = 0; q < NUM_CLS; ++q) {
		if (c->cls_stream[q]) (void)hipStreamDestroy(c->cls_stream[q]);
		if (c->join_ev[q]) (void)hipEventDestroy(c->join_ev[q]);
	}
	if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
	if (c->stream) (void)hipStreamDestroy(c->stream);
	delete c;
}

const char* umans_last_error(const umans_ctx* c) { return c ? c->err.c_str() : "null context"; }

int umans_set_policies(umans_ctx* c, const umans_policy* p, int n) {
	if (!c) return UMANS_EINVAL;
	if (!p || n < 1 || n > kMaxPolicies) return c->fail(UMANS_EINVAL, "need 1..1024 policies");
	std::vector<DevPolicy> v(n);
	float mr = 0.0f;
	for (int i = 0; i < n; ++i) {
		std::string e;
		const int rc = make_dev_policy(p[i], v[i], e);
		if (rc) return c->fail(rc, "policy " + std::to_string(i) + ": " + e);
		mr = std::max(mr, v[i].query_range);
	}
	if (c->n > 0 && mr > c->max_range) return c->fail(UMANS_ESTATE, "a larger interaction range needs a new upload");
	if (c->n > 0 && n < (int)c->hpol.size())
		return c->fail(UMANS_ESTATE, "uploaded agents may use every policy: the table cannot shrink before a new upload");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = c->hip(hipMemcpyAsync(c->dpol, v.data(), n * sizeof(DevPolicy), hipMemcpyHostToDevice, c->stream),
	                      "policy upload");
	if (rc) return rc;
	c->hpol.swap(v);
	c->adj_ready = false;
	c->drop_graph();
	if (c->n == 0) c->max_range = mr;
	else if (int rc2 = update_classes(c)) return rc2;
	return c->hip(hipStreamSynchronize(c->stream), "policy upload");
}

int umans_upload(umans_ctx* c, const umans_agents* a, int n) {
	if (!c) return UMANS_EINVAL;
	if (c->hpol.empty()) return c->fail(UMANS_ESTATE, "set policies before uploading agents");
	if (!a || n < 0 || n > (int)kIndexMask) return c->fail(UMANS_EINVAL, "bad agent count");
	if (n > 0 && (!a->id || !a->pos || !a->goal || !a->radius || !a->pref_speed || !a->max_speed || !a->max_accel || !a->policy))
		return c->fail(UMANS_EINVAL, "missing agent array");
	const int npol = (int)c->hpol.size();
	const int nw = (int)c->worlds.size();
	std::vector<float2> pos(n), vel(n), goal(n), view(n);
	std::vector<float4> prm(n);
	std::vector<int4> meta(n);
	std::vector<int> per_world(nw, 0);
	c->max_range = 0.0f;
	for (const DevPolicy& p : c->hpol) c->max_range = std::max(c->max_range, p.query_range);
	for (int i = 0; i < n; ++i) {
		if (a->policy[i] < 0 || a->policy[i] >= npol)
			return c->fail(UMANS_EINVAL, "agent " + std::to_string(i) + " has an unknown policy id");
		const int w = a->world ? a->world[i] : 0;
		if (w < 0 || w >= nw) return c->fail(UMANS_EINVAL, "agent " + std::to_string(i) + " has an unknown world");
		per_world[w]++;
		pos[i] = make_float2(a->pos[2 * i], a->pos[2 * i + 1]);
		vel[i] = a->vel ? make_float2(a->vel[2 * i], a->vel[2 * i + 1]) : make_float2(0.0f, 0.0f);
		goal[i] = make_float2(a->goal[2 * i], a->goal[2 * i + 1]);
		if (a->viewdir) view[i] = make_float2(a->viewdir[2 * i], a->viewdir[2 * i + 1]);
		else {
			// Agent::setGoal looks at the goal (agent.cpp:120-127)
			const float dx = goal[i].x - pos[i].x, dy = goal[i].y - pos[i].y;
			const float m = std::sqrt(dx * dx + dy * dy);
			view[i] = m > 0.0f ? make_float2(dx / m, dy / m) : make_float2(dx, dy);
		}
		prm[i] = make_float4(a->radius[i], a->pref_speed[i], a->max_speed[i], a->max_accel[i]);
		uint32_t r;
		if (a->rng) r = a->rng[i];
		else {
			// std::default_random_engine(id) = minstd_rand0 seeded with the ID (agent.cpp:38)
			r = (uint32_t)((uint32_t)a->id[i] % 2147483647u);
			if (r == 0) r = 1;
		}
		meta[i] = make_int4(a->id[i], a->policy[i] | (w << kPolicyBits), i, (int)r);
	}
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	if (c->csv_on) {                                  // the logged rows belong to the old agent set
		const int rc0 = csv_flush(c);
		c->csv_on = false;
		c->csv_log.clear();
		if (rc0) return rc0;
	}
	c->drop_graph();
	if (n > c->cap) {
		c->free_agents();
		int rc;
		if ((rc = c->hip(c->X.alloc(n), "alloc state"))) return rc;
		if ((rc = c->hip(c->Y.alloc(n), "alloc state"))) return rc;
		if ((rc = c->hip(dalloc(&c->cell_of, n), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->class_list, n), "alloc grid"))) return rc;
		if (kmax_bucket(c->kcap) > 32) {
			if ((rc = c->hip(dalloc(&c->long_key, (size_t)kLongCap * n), "alloc lists"))) return rc;
			if ((rc = c->hip(dalloc(&c->long_u, (size_t)kLongCap * n), "alloc lists"))) return rc;
		}
		if ((rc = c->hip(dalloc(&c->ovf[0], n), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->ovf[1], n), "alloc grid"))) return rc;
		c->cap = n;
		if ((rc = ensure_huge(c))) return rc;
	}
	c->n = n;
	c->h_policy.assign(a->policy, a->policy + n);
	c->have_record = false;
	// the uploaded agents are the whole list, in row order; nothing is scheduled
	c->row_id.assign(a->id, a->id + n);
	c->row_of.clear();
	c->row_index_built = false;
	c->next_id = 0;
	for (int i = 0; i < n; ++i) c->next_id = std::max(c->next_id, (long long)a->id[i] + 1);
	c->h_remove.assign(n, 0);
	c->n_remove = 0;
	c->remove_dirty = false;
	c->pop_open = false;
	c->src.clear();
	c->fresh.clear();
	c->edits.clear();
	c->pending_recs.clear();
	c->pending = decltype(c->pending)();
	int rc;
	if ((rc = setup_worlds(c, per_world))) return rc;
	if ((rc = update_classes(c))) return rc;
	const State& s = c->X.s;
	if ((rc = c->hip(hipMemcpyAsync(s.pos, pos.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.vel, vel.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.goal, goal.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.viewdir, view.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.prm, prm.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.meta, meta.data(), n * sizeof(int4), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	return c->hip(hipStreamSynchronize(c->stream), "upload");
}

int umans_num_agents(const umans_ctx* c) { return c ? (int)c->row_id.size() : 0; }

int umans_set_record_neighbors(umans_ctx* c, int enable) {
	if (!c) return UMANS_EINVAL;
	if (enable && !c->rec_count) {
		if (c->cap <= 0) return c->fail(UMANS_ESTATE, "upload agents before enabling the neighbour record");
		if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
		const int stride = rec_stride_of(c);
		int rc;
		if ((rc = c->hip(dalloc(&c->rec_count, c->cap), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_id, (size_t)c->cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_d2, (size_t)c->cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_seg, (size_t)c->cap * stride), "alloc record"))) return rc;
		c->rec_stride = stride;
	}
	c->record = enable ? 1 : 0;
	c->drop_graph();
	return UMANS_OK;
}

int umans_set_graph(umans_ctx* c, int enable) {
	if (!c) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = c->hip(hipStreamSynchronize(c->stream), "graph");
	c->use_graph = enable ? 1 : 0;
	c->drop_graph();
	return rc;
}

int umans_step(umans_ctx* c, int nsteps) {
	if (!c) return UMANS_EINVAL;
	if (nsteps < 0) return c->fail(UMANS_EINVAL, "negative step count");
	if (c->hpol.empty()) return c->fail(UMANS_ESTATE, "no policies");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = run_steps(c, nsteps);
	if (rc) return rc;
	if (c->record && nsteps > 0 && c->rec_count) c->have_record = true;
	return UMANS_OK;
}

int umans_autotune(umans_ctx* c, int trial_steps, umans_tuning* out) {
	if (!c) return UMANS_EINVAL;
	if (trial_steps < 1) trial_steps = 3;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc;
	if ((rc = sync_population(c))) return rc;
	if (c->n == 0) return c->fail(UMANS_ESTATE, "upload agents before autotuning");
	std::vector<int> per_world;
	if ((rc = agents_per_world(c, per_world))) return rc;
	if (!c->B.s.pos && (rc = c->hip(c->B.alloc(c->cap), "alloc backup"))) return rc;
	if ((rc = c->hip(copy_state(c->B.s, c->X.s, c->n, c->stream), "backup"))) return rc;
	const int kmax = kmax_bucket(c->kcap);
	const float factors[] = {0.5f, 0.625f, 0.75f, 0.875f, 1.0f, 1.5f};
	const int blocks[] = {64, 128, 256};
	const float f0 = c->cell_factor;
	const int b0 = c->block;
	const int tier0 = c->tier_first;
	const int n_tiers = kmax > 32 ? 4 : 1;   // list choices only matter for long lists
	int best_t = tier0;
	const int rec0 = c->record;
	c->record = 0;
	int cb0[NUM_CLS], best_cb[NUM_CLS] = {};
	for (int q = 0; q < NUM_CLS; ++q) cb0[q] = c->cls_block[q];
	float best_ms = 3.4e38f, best_f = f0;
	int best_b = b0 > 0 ? b0 : default_block(kmax);
	hipEvent_t t0 = nullptr, t1 = nullptr;
	if ((rc = c->hip(hipEventCreate(&t0), "autotune")) || (rc = c->hip(hipEventCreate(&t1), "autotune"))) {
		if (t0) (void)hipEventDestroy(t0);
		return rc;
	}
	// time trial_steps steps of the configuration in c (after a warm-up step)
	auto trial = [&](float& ms) -> int {
		c->drop_graph();
		int r;
		if ((r = setup_worlds(c, per_world))) return r;
		if ((r = c->hip(copy_state(c->X.s, c->B.s, c->n, c->stream), "restore"))) return r;
		if ((r = enqueue_step(c, false))) return r;   // warm-up
		if ((r = c->hip(hipEventRecord(t0, c->stream), "autotune"))) return r;
		for (int s = 0; s < trial_steps && !r; ++s) r = enqueue_step(c, false);
		if (r) return r;
		if ((r = c->hip(hipEventRecord(t1, c->stream), "autotune"))) return r;
		if ((r = c->hip(hipEventSynchronize(t1), "autotune"))) return r;
		ms = 0.0f;
		return c->hip(hipEventElapsedTime(&ms, t0, t1), "autotune");
	};
	// 1. cell width, block size and first list capacity for every class together
	for (int q = 0; q < NUM_CLS; ++q) c->cls_block[q] = 0;
	for (float f : factors) {
		for (int b : blocks)
		for (int ti = 0; ti < n_tiers; ++ti) {
			if (!block_supported(kmax, b)) continue;
			c->cell_factor = f;
			c->block = b;
			c->tier_first = n_tiers > 1 ? ti : tier0;
			float ms;
			if ((rc = trial(ms))) break;
			if (ms < best_ms) { best_ms = ms; best_f = f; best_b = b; best_t = c->tier_first; }
		}
		if (rc) break;
	}
	// 2. then each class's block size in turn, the others kept at their best (the class
	// kernels share the GPU, so one class's best depends on the others')
	if (!rc) {
		c->cell_factor = best_f;
		c->block = best_b;
		c->tier_first = best_t;
		// a fresh time of the stage-1 choice (its best trial is the luckiest of many), and a
		// per-class block kept only if it is faster by more than the trials' noise (1 %)
		for (int p = 0; p < NUM_CLS; ++p) c->cls_block[p] = best_cb[p];
		float base_ms;
		if (!(rc = trial(base_ms))) best_ms = base_ms;
		for (int q = 0; q < NUM_CLS && !rc; ++q) {
			if (q == CLS_COOP || c->cls_count[q] == 0) continue;
			for (int b : blocks) {
				if (b == best_b || !block_supported(kmax, b)) continue;
				for (int p = 0; p < NUM_CLS; ++p) c->cls_block[p] = best_cb[p];
				c->cls_block[q] = b;
				float ms;
				if ((rc = trial(ms))) break;
				if (ms < 0.99f * best_ms) { best_ms = ms; for (int p = 0; p < NUM_CLS; ++p) best_cb[p] = c->cls_block[p]; }
			}
		}
	}
	(void)hipEventDestroy(t0);
	(void)hipEventDestroy(t1);
	// restore the configuration chosen (or the previous one on failure) and the state
	c->cell_factor = rc ? f0 : best_f;
	c->block = rc ? b0 : best_b;
	c->tier_first = rc ? tier0 : best_t;
	for (int q = 0; q < NUM_CLS; ++q) c->cls_block[q] = rc ? cb0[q] : best_cb[q];
	c->record = rec0;
	c->drop_graph();
	int rc2 = setup_worlds(c, per_world);
	if (!rc2) rc2 = c->hip(copy_state(c->X.s, c->B.s, c->n, c->stream), "restore");
	if (!rc2) rc2 = c->hip(hipStreamSynchronize(c->stream), "autotune");
	if (rc) return rc;
	if (rc2) return rc2;
	if (out) {
		out->cell_factor = c->cell_factor;
		out->block_size = c->block;
		out->ms_per_step = best_ms / trial_steps;
		out->list_capacity = c->tier_first == 0 ? 32 : (c->tier_first == 1 ? 64 : (c->tier_first == 2 ? 128 : 1));
		for (int q = 0; q < NUM_CLS; ++q) out->class_block[q] = c->cls_block[q];
	}
	return UMANS_OK;
}

int umans_set_tuning(umans_ctx* c, const umans_tuning* t) {
	if (!c || !t) return UMANS_EINVAL;
	if (!(t->cell_factor > 0.0f) || (t->block_size != 0 && !block_supported(kmax_bucket(c->kcap), t->block_size)))
		return c->fail(UMANS_EINVAL, "unsupported tuning");
	if (t->list_capacity != 0 && t->list_capacity != 1 && t->list_capacity != 32 && t->list_capacity != 64 &&
	    t->list_capacity != 128)
		return c->fail(UMANS_EINVAL, "list_capacity must be 0, 1, 32, 64 or 128");
	for (int q = 0; q < NUM_CLS; ++q)
		if (t->class_block[q] != 0 && !block_supported(kmax_bucket(c->kcap), t->class_block[q]))
			return c->fail(UMANS_EINVAL, "unsupported class block size");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc = c->hip(hipStreamSynchronize(c->stream), "tuning");
	if (rc) return rc;
	c->cell_factor = t->cell_factor;
	c->block = t->block_size;
	for (int q = 0; q < NUM_CLS; ++q) c->cls_block[q] = t->class_block[q];
	if (t->list_capacity)
		c->tier_first = t->list_capacity == 1 ? 3 : (t->list_capacity == 32 ? 0 : (t->list_capacity == 64 ? 1 : 2));
	c->drop_graph();
	if (c->n > 0) {
		std::vector<int> per_world;
		if ((rc = agents_per_world(c, per_world))) return rc;
		if ((rc = setup_worlds(c, per_world))) return rc;
		rc = c->hip(hipStreamSynchronize(c->stream), "tuning");
	}
	return rc;
}

int umans_csv_start(umans_ctx* c, const char* dirname, int flush_immediately) {
	if (!c || !dirname || !*dirname) return UMANS_EINVAL;
	if (c->worlds.size() != 1) return c->fail(UMANS_EINVAL, "CSV output is per world: use a single-world context");
	if (c->csv_on) {
		if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
		const int rc0 = csv_flush(c);
		if (rc0) return rc0;
	}
	std::string dir = dirname;
	if (dir.back() != '/') dir += '/';                // CSVWriter::SetOutputDirectory (csvwriter.cpp:37-48)
	for (size_t i = 1; i <= dir.size(); ++i)
		if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
	struct stat sb;
	if (stat(dir.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode))
		return c->fail(UMANS_EINVAL, "could not set CSV output directory to " + dir);
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc = sync_population(c);
	if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "csv");
	if (rc) return rc;
	if (c->csv_cap < c->n && (rc = csv_alloc(c, c->n))) return rc;
	c->csv_log.clear();
	c->csv_slot_id.clear();
	c->csv_slot_of_id.clear();
	csv_rows_changed(c);
	c->csv_dir = dir;
	c->csv_flush_now = flush_immediately != 0;
	c->csv_on = true;
	return UMANS_OK;
}

int umans_csv_flush(umans_ctx* c) {
	if (!c) return UMANS_EINVAL;
	if (!c->csv_on) return UMANS_OK;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	return csv_flush(c);
}

int umans_csv_stop(umans_ctx* c) {
	if (!c) return UMANS_EINVAL;
	if (!c->csv_on) return UMANS_OK;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = csv_flush(c);
	c->csv_on = false;
	c->csv_log.clear();
	return rc;
}

int umans_set_profiling(umans_ctx* c, int enable) {
	if (!c) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc = c->hip(hipStreamSynchronize(c->stream), "profiling");
	if (rc) return rc;
	c->ev_used = 0;
	c->prof_steps = 0;
	c->prof_grid_ms = c->prof_step_ms = 0.0;
	c->profiling = enable ? 1 : 0;
	return UMANS_OK;
}

int umans_get_profile(umans_ctx* c, umans_profile* out) {
	if (!c || !out) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = c->drain_profile();
	if (rc) return rc;
	out->steps = c->prof_steps;
	out->grid_ms = c->prof_grid_ms;
	out->step_ms = c->prof_step_ms;
	return UMANS_OK;
}

int umans_synchronize(umans_ctx* c) {
	if (!c) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	if (const int rc = c->hip(hipStreamSynchronize(c->stream), "synchronize")) return rc;
	return take_flags(c, kFlagListTooLong);
}

int umans_get_stats(umans_ctx* c, umans_stats* out) {
	if (!c || !out) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	std::vector<unsigned long long> hc((size_t)kStatCopies * kStatLine);
	int rc = c->hip(hipMemcpyAsync(hc.data(), c->dstats, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
	                               c->stream), "stats");
	if (rc) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "stats"))) return rc;
	unsigned long long h[3] = {0, 0, 0};
	for (int cp = 0; cp < kStatCopies; ++cp) {
		const unsigned long long* s = hc.data() + (size_t)cp * kStatLine;
		h[0] += s[0];
		h[1] = std::max(h[1], s[1]);
		h[2] += s[2];
	}
	out->steps = c->steps;
	out->overflow_agents = (int64_t)h[0];
	out->max_list = (int64_t)h[1];
	out->sim_time = c->sim_time;
	out->obstacle_overflow_agents = (int64_t)h[2];
	return take_flags(c, kFlagListTooLong);
}

int umans_download(umans_ctx* c, umans_agents* a) {
	if (!c || !a) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	if (int rc0 = sync_population(c)) return rc0;
	const int n = c->n;
	std::vector<float2> pos(n), vel(n), view(n), goal(n), pref(n);
	std::vector<float4> prm(n);
	std::vector<int4> meta(n);
	const State& s = c->X.s;
	int rc;
	if ((rc = c->hip(hipMemcpyAsync(pos.data(), s.pos, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(vel.data(), s.vel, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(view.data(), s.viewdir, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(goal.data(), s.goal, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(pref.data(), s.pref, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(prm.data(), s.prm, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(meta.data(), s.meta, n * sizeof(int4), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "download"))) return rc;
	for (int k = 0; k < n; ++k) {
		const int i = meta[k].z;   // upload row
		if (i < 0 || i >= n) return c->fail(UMANS_EHIP, "corrupt agent row");
		if (a->id) a->id[i] = meta[k].x;
		if (a->policy) a->policy[i] = meta[k].y & kPolicyMask;
		if (a->world) a->world[i] = meta[k].y >> kPolicyBits;
		if (a->rng) a->rng[i] = (uint32_t)meta[k].w;
		if (a->pos) { a->pos[2 * i] = pos[k].x; a->pos[2 * i + 1] = pos[k].y; }
		if (a->vel) { a->vel[2 * i] = vel[k].x; a->vel[2 * i + 1] = vel[k].y; }
		if (a->viewdir) { a->viewdir[2 * i] = view[k].x; a->viewdir[2 * i + 1] = view[k].y; }
		if (a->goal) { a->goal[2 * i] = goal[k].x; a->goal[2 * i + 1] = goal[k].y; }
		if (a->pref_vel) { a->pref_vel[2 * i] = pref[k].x; a->pref_vel[2 * i + 1] = pref[k].y; }
		if (a->radius) a->radius[i] = prm[k].x;
		if (a->pref_speed) a->pref_speed[i] = prm[k].y;
		if (a->max_speed) a->max_speed[i] = prm[k].z;
		if (a->max_accel) a->max_accel[i] = prm[k].w;
	}
	return take_flags(c, kFlagListTooLong);
}

int umans_dump_neighbors(umans_ctx* c, int stride, int32_t* count, int32_t* ids, double* d2, int32_t* seg) {
	if (!c || !count || stride < 0 || (stride > 0 && (!ids || !d2 || !seg))) return UMANS_EINVAL;
	if (!c->have_record) return c->fail(UMANS_ESTATE, "no neighbour record: enable it and run a step");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int n = c->n, rs = c->rec_stride;
	std::vector<int> hc(n), hid((size_t)n * rs), hseg((size_t)n * rs);
	std::vector<double> hd((size_t)n * rs);
	std::vector<int4> meta(n);
	int rc;
	if ((rc = c->hip(hipMemcpyAsync(hc.data(), c->rec_count, n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "dump"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(hid.data(), c->rec_id, hid.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream), "dump"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(hd.data(), c->rec_d2, hd.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "dump"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(hseg.data(), c->rec_seg, hseg.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream), "dump"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(meta.data(), c->X.s.meta, n * sizeof(int4), hipMemcpyDeviceToHost, c->stream), "dump"))) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "dump"))) return rc;
	for (int k = 0; k < n; ++k) {
		const int i = meta[k].z;
		count[i] = hc[k];
		const int m = std::min(std::min(hc[k], rs), stride);
		for (int e = 0; e < m; ++e) {
			ids[(size_t)i * stride + e] = hid[(size_t)k * rs + e];
			d2[(size_t)i * stride + e] = hd[(size_t)k * rs + e];
			seg[(size_t)i * stride + e] = hseg[(size_t)k * rs + e];
		}
	}
	return UMANS_OK;
}

int umans_adjoint_forward(umans_ctx* c, const umans_adjoint_desc* d, const float* targets, double* loss) {
	if (!c || !d || !targets) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	if (int rc0 = sync_population(c)) return rc0;
	if (c->n == 0) return c->fail(UMANS_ESTATE, "upload agents first");
	if (!c->pending.empty() || c->n_remove > 0)
		return c->fail(UMANS_ESTATE, "the adjoint needs a fixed population (no scheduled agents, no remove_at_goal)");
	if (c->n_edges > 0) return c->fail(UMANS_ESTATE, "the adjoint does not differentiate obstacle forces");
	if (d->nsteps < 1 || d->checkpoint_every < 1) return c->fail(UMANS_EINVAL, "bad step or checkpoint count");
	if (d->param_kind != UMANS_COST_SOCIAL_FORCES && d->param_kind != UMANS_COST_POWER_LAW)
		return c->fail(UMANS_EINVAL, "parameters are differentiated for SocialForcesAvoidance or PowerLaw");
	for (size_t i = 0; i < c->hpol.size(); ++i)
		if (!adjoint_policy_ok(c->hpol[i]))
			return c->fail(UMANS_EINVAL, "the adjoint needs gradient-method policies of force models (policy " +
			                                 std::to_string(i) + ")");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	c->drop_graph();
	c->free_adjoint();
	const int n = c->n, T = d->nsteps, ck = d->checkpoint_every;
	const int nseg = ceil_div(T, ck);
	const int kmax = kmax_bucket(c->kcap);
	int rc;
	c->ckpts.resize(nseg);
	for (DeviceState& s : c->ckpts)
		if ((rc = c->hip(s.alloc(n), "alloc checkpoints"))) return rc;
	c->seg_states.resize(std::min(ck, T));
	for (DeviceState& s : c->seg_states)
		if ((rc = c->hip(s.alloc(n), "alloc segment"))) return rc;
	const int segn = (int)c->seg_states.size();
	if ((rc = c->hip(dalloc(&c->targets, (size_t)T * n), "alloc targets"))) return rc;
	if ((rc = c->hip(dalloc(&c->seg_nb, (size_t)segn * n * kmax), "alloc lists"))) return rc;
	if ((rc = c->hip(dalloc(&c->seg_cnt, (size_t)segn * n), "alloc lists"))) return rc;
	if ((rc = c->hip(dalloc(&c->adj_buf, (size_t)5 * n), "alloc adjoint"))) return rc;
	if ((rc = c->hip(dalloc(&c->adj_out, (size_t)n), "alloc adjoint"))) return rc;
	if ((rc = c->hip(dalloc(&c->adj_clamps, (size_t)T * n), "alloc adjoint"))) return rc;
	if ((rc = c->hip(dalloc(&c->adj_params, (size_t)kStatCopies * adj_param_stride((int)c->hpol.size())), "alloc adjoint")))
		return rc;
	if ((rc = c->hip(dalloc(&c->dloss, 1), "alloc loss"))) return rc;
	// host or device memory (unified addressing)
	if ((rc = c->hip(hipMemcpyAsync(c->targets, targets, (size_t)T * n * sizeof(float2), hipMemcpyDefault,
	                                c->stream), "targets"))) return rc;
	fill_f(c, c->dloss, 2);
	for (int t = 0; t < T; ++t) {
		if (t % ck == 0 && (rc = c->hip(copy_state(c->ckpts[t / ck].s, c->X.s, n, c->stream), "checkpoint"))) return rc;
		if ((rc = enqueue_step(c, false, nullptr, nullptr, nullptr, t + 1 < T))) return rc;
		launch_loss(c, c->targets + (size_t)t * n, c->dloss, nullptr, 0.0);
		c->steps += 1;
		c->sim_time += (double)c->worlds[0].delta_time;
	}
	double h = 0.0;
	if ((rc = c->hip(hipMemcpyAsync(&h, c->dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream), "loss"))) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "loss"))) return rc;
	c->last_loss = h / ((double)T * (double)n);
	c->adj_steps = T;
	c->adj_ckpt = ck;
	c->adj_n = n;
	c->adj_param_kind = d->param_kind;
	c->adj_ready = true;
	if (loss) *loss = c->last_loss;
	return UMANS_OK;
}

int umans_adjoint_backward(umans_ctx* c, float* d_params, float* d_pos0, float* d_vel0) {
	if (!c) return UMANS_EINVAL;
	if (!c->adj_ready || c->adj_n != c->n) return c->fail(UMANS_ESTATE, "run umans_adjoint_forward first");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int n = c->n, T = c->adj_steps, ck = c->adj_ckpt;
	const int nseg = ceil_div(T, ck);
	const int kmax = kmax_bucket(c->kcap);
	const int npol = (int)c->hpol.size();
	const double scale = 1.0 / ((double)T * (double)n);
	int rc;
	if (!c->B.s.pos && (rc = c->hip(c->B.alloc(c->cap), "alloc backup"))) return rc;
	if ((rc = c->hip(copy_state(c->B.s, c->X.s, n, c->stream), "backup"))) return rc;
	double2* ap = c->adj_buf;
	double2* av = c->adj_buf + n;
	double2* ap2 = c->adj_buf + 2 * n;
	double2* av2 = c->adj_buf + 3 * n;
	double2* apref = c->adj_buf + 4 * n;
	// doubles are cleared as zero bits: 4 floats per double2
	fill_f(c, c->adj_buf, 4LL * 5 * n);
	const int pstride = adj_param_stride(npol);
	fill_f(c, c->adj_params, 2LL * kStatCopies * pstride);     // doubles: zero bits
	for (int seg = nseg - 1; seg >= 0; --seg) {
		const int t0 = seg * ck, t1 = std::min(t0 + ck, T);
		if ((rc = c->hip(copy_state(c->X.s, c->ckpts[seg].s, n, c->stream), "restore checkpoint"))) return rc;
		for (int t = t0; t < t1; ++t) {
			const int s = t - t0;
			if ((rc = enqueue_step(c, false, &c->seg_states[s].s, c->seg_nb + (size_t)s * n * kmax,
			                       c->seg_cnt + (size_t)s * n, t + 1 < t1))) return rc;
		}
		// the loss gradient of the last step's output; those of the earlier steps are
		// added by k_adj_link from the next step's input state
		if (seg == nseg - 1) launch_loss(c, c->targets + (size_t)(T - 1) * n, nullptr, ap, scale);
		for (int t = t1 - 1; t >= t0; --t) {
			const int s = t - t0;
			AdjArgs A{};
			A.in = c->seg_states[s].s;
			A.nb_count = c->seg_cnt + (size_t)s * n;
			A.nb = c->seg_nb + (size_t)s * n * kmax;
			A.stride = kmax;
			A.pol = c->dpol;
			A.worlds = c->dworlds;
			A.n = n;
			A.adj_pos_out = ap;
			A.adj_vel_out = av;
			A.adj_pos_in = ap2;
			A.adj_vel_in = av2;
			A.adj_pref = apref;
			A.adj_params = c->adj_params;
			A.param_stride = pstride;
			A.param_kind = c->adj_param_kind;
			A.clamps = c->adj_clamps + (size_t)t * n;
			hipLaunchKernelGGL(k_step_adjoint, dim3(ceil_div(n, 4)), dim3(256), 0, c->stream, A);
			hipLaunchKernelGGL(k_adj_link, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, c->seg_states[s].s, n,
			                   c->dworlds, apref, ap2, t > 0 ? c->targets + (size_t)(t - 1) * n : nullptr, scale,
			                   ap, av);
			std::swap(ap, ap2);
			std::swap(av, av2);
		}
	}
	if ((rc = c->hip(hipGetLastError(), "adjoint launch"))) return rc;
	std::vector<double> hc((size_t)kStatCopies * pstride);
	if ((rc = c->hip(hipMemcpyAsync(hc.data(), c->adj_params, hc.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "adjoint"))) return rc;
	// handed out in float, to host or device memory (unified addressing); the stream
	// orders the staging buffer's second use after the first copy
	const double2* outs[2] = {ap, av};
	float* dsts[2] = {d_pos0, d_vel0};
	for (int q = 0; q < 2; ++q) {
		if (!dsts[q]) continue;
		hipLaunchKernelGGL(k_to_float2, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, outs[q], c->adj_out, n);
		if ((rc = c->hip(hipMemcpyAsync(dsts[q], c->adj_out, (size_t)n * sizeof(float2), hipMemcpyDefault, c->stream),
		                 "adjoint"))) return rc;
	}
	if ((rc = c->hip(copy_state(c->X.s, c->B.s, n, c->stream), "restore"))) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "adjoint"))) return rc;
	if (d_params) {
		for (int i = 0; i < 2 * npol; ++i) {
			double sum = 0.0;
			for (int cp = 0; cp < kStatCopies; ++cp) sum += hc[(size_t)cp * pstride + i];
			d_params[i] = (float)sum;
		}
	}
	c->adj_clamps_ready = true;
	return take_flags(c, kFlagListTooLong | kFlagRawListTooLong);
}

int umans_adjoint_clamps(umans_ctx* c, uint8_t* flags) {
	if (!c || !flags) return UMANS_EINVAL;
	if (!c->adj_ready || !c->adj_clamps_ready || c->adj_n != c->n)
		return c->fail(UMANS_ESTATE, "run umans_adjoint_forward and umans_adjoint_backward first");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	const int rc = c->hip(hipMemcpyAsync(flags, c->adj_clamps, (size_t)c->adj_steps * c->n, hipMemcpyDefault, c->stream),
	                      "adjoint clamps");
	if (rc) return rc;
	return c->hip(hipStreamSynchronize(c->stream), "adjoint clamps");
}

} // extern "C"

// ------------------------------------------------------------- population API
namespace {

int single_world(umans_ctx* c) {
	if (c->worlds.size() != 1) return c->fail(UMANS_EINVAL, "population changes need a single-world context");
	return UMANS_OK;
}

// The list row of an agent ID, or -1.
int find_row(umans_ctx* c, int32_t id) {
	build_row_index(c);
	const auto it = c->row_of.find(id);
	return it == c->row_of.end() ? -1 : it->second;
}

// Agent::setGoal's look direction (agent.cpp:120-127; Vector2D::normalize, vector2D.h:55-63).
bool look_at_goal(float2 pos, float2 goal, float2& view) {
	if (goal.x == pos.x && goal.y == pos.y) return false;
	const float dx = goal.x - pos.x, dy = goal.y - pos.y;
	const float mag = std::sqrt(dx * dx + dy * dy);
	view = mag > 0.0f ? make_float2(dx / mag, dy / mag) : make_float2(dx, dy);
	return true;
}

} // namespace

extern "C" {

int umans_add_agents(umans_ctx* c, const umans_agents* a, int n, const float* start_time,
                     const uint8_t* remove_at_goal, int flags, int32_t* out_ids) {
	if (!c) return UMANS_EINVAL;
	if (int rc = single_world(c)) return rc;
	if (c->hpol.empty()) return c->fail(UMANS_ESTATE, "set policies before adding agents");
	if (n < 0) return c->fail(UMANS_EINVAL, "bad agent count");
	if (n > 0 && (!a || !a->pos || !a->radius || !a->pref_speed || !a->max_speed || !a->max_accel || !a->policy))
		return c->fail(UMANS_EINVAL, "missing agent array");
	const int npol = (int)c->hpol.size();
	for (int i = 0; i < n; ++i)
		if (a->policy[i] < 0 || a->policy[i] >= npol)
			return c->fail(UMANS_EINVAL, "agent " + std::to_string(i) + " has an unknown policy id");
	if ((long long)c->row_id.size() + (long long)c->pending.size() + n > (long long)kIndexMask)
		return c->fail(UMANS_EINVAL, "too many agents");
	build_row_index(c);
	for (int i = 0; i < n; ++i) {
		// WorldBase::AddAgent (worldBase.cpp:161-187): a desired ID is used when no agent
		// in the list has it, otherwise the next unused ID
		const long long desired = a->id ? (long long)a->id[i] : -1;
		long long id;
		if (flags & UMANS_ADD_KEEP_IDS) {
			if (desired < 0) return c->fail(UMANS_EINVAL, "UMANS_ADD_KEEP_IDS needs an ID for every agent");
			id = desired;
		} else {
			id = (desired >= 0 && !c->row_of.count((int32_t)desired)) ? desired : c->next_id;
		}
		if (id > 0x7fffffffLL) return c->fail(UMANS_EINVAL, "agent IDs exhausted");
		if (id >= c->next_id) c->next_id = id + 1;
		umans_ctx::AgentRec r;
		r.pos = make_float2(a->pos[2 * i], a->pos[2 * i + 1]);
		r.vel = a->vel ? make_float2(a->vel[2 * i], a->vel[2 * i + 1]) : make_float2(0.0f, 0.0f);
		// without a goal the agent stays where it is (APIFunctions.cpp:163-164)
		r.goal = a->goal ? make_float2(a->goal[2 * i], a->goal[2 * i + 1]) : r.pos;
		r.view = make_float2(0.0f, 0.0f);
		if (a->viewdir) r.view = make_float2(a->viewdir[2 * i], a->viewdir[2 * i + 1]);
		else look_at_goal(r.pos, r.goal, r.view);
		r.prm = make_float4(a->radius[i], a->pref_speed[i], a->max_speed[i], a->max_accel[i]);
		uint32_t seed;
		if (a->rng) seed = a->rng[i];
		else {
			seed = (uint32_t)((uint32_t)id % 2147483647u);   // minstd_rand0 seeded with the ID (agent.cpp:38)
			if (seed == 0) seed = 1;
		}
		r.meta = make_int4((int)id, a->policy[i], -1, (int)seed);
		const unsigned char rm = remove_at_goal && remove_at_goal[i] ? 1 : 0;
		const float t = start_time ? start_time[i] : 0.0f;
		if ((double)t <= c->sim_time) {
			list_append(c, r, rm);
		} else {
			c->pending.push({(int)c->pending_recs.size(), t});
			c->pending_recs.push_back({r, rm});
		}
		if (out_ids) out_ids[i] = (int32_t)id;
	}
	return UMANS_OK;
}

int umans_remove_agent(umans_ctx* c, int32_t id) {
	if (!c) return UMANS_EINVAL;
	if (int rc = single_world(c)) return rc;
	const int r = find_row(c, id);
	if (r < 0) return c->fail(UMANS_ENOTFOUND, "no agent with ID " + std::to_string(id));
	list_remove_row(c, r);
	return UMANS_OK;
}

int umans_set_agent_goal(umans_ctx* c, int32_t id, float x, float y) {
	if (!c) return UMANS_EINVAL;
	if (int rc = single_world(c)) return rc;
	const int r = find_row(c, id);
	if (r < 0) return c->fail(UMANS_ENOTFOUND, "no agent with ID " + std::to_string(id));
	open_edits(c);
	const int h = c->src[r];
	const float2 g = make_float2(x, y);
	if (h < 0) {
		umans_ctx::AgentRec& rec = c->fresh[-1 - h];
		rec.goal = g;
		look_at_goal(rec.pos, g, rec.view);
		return UMANS_OK;
	}
	auto it = c->edits.find(h);
	if (it != c->edits.end() && (it->second.mask & EDIT_VIEW_FROM_GOAL)) {
		// a second goal before the device has the first one's look direction
		if (int rc = apply_device_edits(c)) return rc;
		it = c->edits.end();
	}
	if (it == c->edits.end()) {
		PopEdit e{};
		e.row = h;
		it = c->edits.emplace(h, e).first;
	}
	PopEdit& e = it->second;
	e.mask |= EDIT_GOAL;
	e.goal = g;
	if (e.mask & EDIT_POS) {
		if (look_at_goal(e.pos, g, e.view)) e.mask |= EDIT_VIEW;
	} else {
		e.mask |= EDIT_VIEW_FROM_GOAL;
	}
	return UMANS_OK;
}

int umans_set_agent_states(umans_ctx* c, int n, const int32_t* ids, const float* pos, const float* vel,
                           const float* viewdir) {
	if (!c) return UMANS_EINVAL;
	if (int rc = single_world(c)) return rc;
	if (n < 0 || (n > 0 && (!ids || !pos || !vel || !viewdir))) return c->fail(UMANS_EINVAL, "bad agent arrays");
	for (int i = 0; i < n; ++i) {
		// SetAgentPositions (APIFunctions.cpp:117-137): unknown IDs are skipped
		const int r = find_row(c, ids[i]);
		if (r < 0) continue;
		open_edits(c);
		const float2 p = make_float2(pos[2 * i], pos[2 * i + 1]);
		const float2 v = make_float2(vel[2 * i], vel[2 * i + 1]);
		const float2 w = make_float2(viewdir[2 * i], viewdir[2 * i + 1]);
		const int h = c->src[r];
		if (h < 0) {
			umans_ctx::AgentRec& rec = c->fresh[-1 - h];
			rec.pos = p; rec.vel = v; rec.view = w;
			continue;
		}
		auto it = c->edits.find(h);
		if (it == c->edits.end()) {
			PopEdit e{};
			e.row = h;
			it = c->edits.emplace(h, e).first;
		}
		PopEdit& e = it->second;
		e.mask = (e.mask | EDIT_POS | EDIT_VEL | EDIT_VIEW) & ~EDIT_VIEW_FROM_GOAL;
		e.pos = p; e.vel = v; e.view = w;
	}
	return UMANS_OK;
}

int umans_set_remove_at_goal(umans_ctx* c, int n, const int32_t* ids, const uint8_t* flags) {
	if (!c) return UMANS_EINVAL;
	if (int rc = single_world(c)) return rc;
	if (n < 0 || (n > 0 && (!ids || !flags))) return c->fail(UMANS_EINVAL, "bad arrays");
	for (int i = 0; i < n; ++i) {
		const int r = find_row(c, ids[i]);
		if (r < 0) return c->fail(UMANS_ENOTFOUND, "no agent with ID " + std::to_string(ids[i]));
		c->h_remove[r] = flags[i] ? 1 : 0;
	}
	c->remove_dirty = true;
	return UMANS_OK;
}

int umans_num_scheduled(const umans_ctx* c) { return c ? (int)c->pending.size() : 0; }

int umans_set_obstacles(umans_ctx* c, int n_polygons, const int32_t* counts, const float* points) {
	if (!c) return UMANS_EINVAL;
	if (c->worlds.size() != 1) return c->fail(UMANS_EINVAL, "obstacles need a single-world context");
	if (n_polygons < 0 || (n_polygons > 0 && (!counts || !points))) return c->fail(UMANS_EINVAL, "bad polygon arrays");
	// Polygon2D::Polygon2D (Polygon2D.cpp:27-38): edge i joins vertex i and vertex (i + 1) mod n
	std::vector<float4> edges;
	long long o = 0;
	for (int p = 0; p < n_polygons; ++p) {
		const int m = counts[p];
		if (m < 0) return c->fail(UMANS_EINVAL, "negative vertex count");
		for (int i = 0; i < m; ++i) {
			const float* a = points + 2 * (o + i);
			const float* b = points + 2 * (o + (i + 1) % m);
			edges.push_back(make_float4(a[0], a[1], b[0], b[1]));
		}
		o += m;
	}
	if (edges.size() > (size_t)(1 << 24)) return c->fail(UMANS_EINVAL, "too many obstacle edges");
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc = c->hip(hipStreamSynchronize(c->stream), "obstacles");
	if (rc) return rc;
	(void)hipFree(c->d_edges);
	c->d_edges = nullptr;
	c->n_edges = 0;
	if (!edges.empty()) {
		if ((rc = c->hip(dalloc(&c->d_edges, edges.size()), "alloc obstacles"))) return rc;
		if ((rc = c->hip(hipMemcpy(c->d_edges, edges.data(), edges.size() * sizeof(float4), hipMemcpyHostToDevice),
		                 "obstacles"))) return rc;
		c->n_edges = (int)edges.size();
	}
	c->drop_graph();
	c->adj_ready = false;
	if ((rc = ensure_huge(c))) return rc;
	return update_classes(c);
}

int umans_download_records(umans_ctx* c, umans_agent_record* out, int capacity) {
	static_assert(sizeof(umans_agent_record) == sizeof(AgentRecord), "record layout");
	if (!c || (!out && capacity > 0)) return UMANS_EINVAL;
	if (const int rc_dev = c->hip(hipSetDevice(c->device), "set device")) return rc_dev;
	int rc = sync_population(c);
	if (rc) return rc;
	const int n = c->n;
	if (n > capacity) return c->fail(UMANS_EINVAL, "record buffer holds " + std::to_string(capacity) +
	                                 " agents, the list has " + std::to_string(n));
	if (n == 0) return UMANS_OK;
	AgentRecord* d = nullptr;
	if ((rc = c->hip(dalloc(&d, n), "alloc records"))) return rc;
	hipLaunchKernelGGL(k_pack_records, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, c->X.s, n, d);
	rc = c->hip(hipMemcpyAsync(out, d, n * sizeof(AgentRecord), hipMemcpyDeviceToHost, c->stream), "records");
	if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "records");
	(void)hipFree(d);
	return rc;
}

} // extern "C"

// ------------------------------------------------- steps without device allocation
// umans_step_noalloc runs the kernels of umans_step on caller-owned device arrays. Every
// buffer beyond its inputs and outputs (the packed and the cell-sorted state, the grid's
// counters, scan and class lists, the policy table, the step counters) lives in one
// file-scope device array, `scratch`, of fixed size; each GPU the library runs on holds
// its own copy until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the device address
// hipGetSymbolAddress returns, looked up once per device.
namespace {

constexpr int kNoAllocMaxPolicies = 16;
constexpr int kNoAllocMaxCells = 4 * UMANS_NOALLOC_MAX_AGENTS > 4096 ? 4 * UMANS_NOALLOC_MAX_AGENTS : 4096;
constexpr int kNoAllocMaxDevices = 64;

constexpr size_t na_up(size_t x) { return (x + 255) & ~(size_t)255; }
// Byte offsets of the sections of `scratch`, each on a 256-byte boundary.
struct NaLayout {
	static constexpr size_t N = UMANS_NOALLOC_MAX_AGENTS;
	static constexpr size_t KEYS = (size_t)NUM_CLS * kNoAllocMaxCells + 1;      // (class slot, cell) counters + total
	static constexpr size_t TILES = (KEYS + kScanTile - 1) / kScanTile;
	static constexpr size_t pol = 0;
	static constexpr size_t world = na_up(pol + kNoAllocMaxPolicies * sizeof(DevPolicy));
	static constexpr size_t grid = na_up(world + sizeof(WorldInfo));
	static constexpr size_t bbox = na_up(grid + sizeof(GridInfo));
	static constexpr size_t stats = na_up(bbox + (size_t)kBoxCopies * kBoxLine * sizeof(int));
	static constexpr size_t cls_slot = na_up(stats + (size_t)kStatCopies * kStatLine * sizeof(unsigned long long));
	static constexpr size_t ranges = na_up(cls_slot + NUM_CLS * sizeof(int));
	static constexpr size_t cell_count = na_up(ranges + 2 * NUM_CLS * sizeof(unsigned));
	static constexpr size_t cell_start = na_up(cell_count + KEYS * sizeof(unsigned));
	static constexpr size_t cell_first = na_up(cell_start + KEYS * sizeof(unsigned));
	static constexpr size_t scan = na_up(cell_first + (kNoAllocMaxCells + 1) * sizeof(unsigned));
	static constexpr size_t cell_of = na_up(scan + (TILES + 1) * sizeof(unsigned));
	static constexpr size_t class_list = na_up(cell_of + N * sizeof(unsigned));
	static constexpr size_t state0 = na_up(class_list + N * sizeof(int));
	// a State: pos, vel, goal, viewdir, pref (float2), prm (float4), meta (int4)
	static constexpr size_t state_bytes = N * (5 * sizeof(float2) + sizeof(float4) + sizeof(int4));
	static constexpr size_t state1 = na_up(state0 + state_bytes);
	static constexpr size_t total = na_up(state1 + state_bytes);
};

__device__ __align__(256) unsigned char scratch[NaLayout::total];

// scratch's device address on each device (null until looked up)
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_scratch_mu;

State na_state(unsigned char* base) {
	const size_t N = NaLayout::N;
	State s;
	s.pos = (float2*)base;
	s.vel = s.pos + N;
	s.goal = s.vel + N;
	s.viewdir = s.goal + N;
	s.pref = s.viewdir + N;
	s.prm = (float4*)(s.pref + N);
	s.meta = (int4*)(s.prm + N);
	return s;
}

// Caller arrays (row order) -> packed state: what umans_upload builds on the host.
__global__ __launch_bounds__(256) void k_noalloc_pack(umans_agents a, int n, int n_policies, State x) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	x.pos[i] = make_float2(a.pos[2 * i], a.pos[2 * i + 1]);
	x.vel[i] = make_float2(a.vel[2 * i], a.vel[2 * i + 1]);
	x.goal[i] = make_float2(a.goal[2 * i], a.goal[2 * i + 1]);
	x.viewdir[i] = make_float2(a.viewdir[2 * i], a.viewdir[2 * i + 1]);
	x.prm[i] = make_float4(a.radius[i], a.pref_speed[i], a.max_speed[i], a.max_accel[i]);
	int p = a.policy[i];
	if (p < 0 || p >= n_policies) p = 0;       // out-of-range indices run policy 0 (documented)
	uint32_t r;
	if (a.rng) r = a.rng[i];
	else {
		// std::default_random_engine(id) = minstd_rand0 seeded with the ID (agent.cpp:38)
		r = (uint32_t)a.id[i] % 2147483647u;
		if (r == 0) r = 1;
	}
	x.meta[i] = make_int4(a.id[i], p, i, (int)r);
}

// Packed state (sorted order) -> caller arrays (row order): pos, vel, viewdir, RNG state and
// the preferred velocity of the last step.
__global__ __launch_bounds__(256) void k_noalloc_unpack(State x, int n, umans_agents a) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int4 m = x.meta[k];
	const int i = m.z;
	const float2 p = x.pos[k], v = x.vel[k], w = x.viewdir[k];
	a.pos[2 * i] = p.x; a.pos[2 * i + 1] = p.y;
	a.vel[2 * i] = v.x; a.vel[2 * i + 1] = v.y;
	a.viewdir[2 * i] = w.x; a.viewdir[2 * i + 1] = w.y;
	if (a.rng) a.rng[i] = (uint32_t)m.w;
	if (a.pref_vel) {
		const float2 q = x.pref[k];
		a.pref_vel[2 * i] = q.x; a.pref_vel[2 * i + 1] = q.y;
	}
}

// First list entry and agent count of each class slot: the class-major scan's start of
// the slot and its difference to the next slot's (the scan has nslots * slot_cells + 1
// entries, the last one the total).
__global__ void k_class_ranges(const unsigned* class_start, int slot_cells, int nslots, unsigned* ranges) {
	const int q = threadIdx.x;
	if (q >= nslots) return;
	const unsigned b = class_start[(size_t)q * slot_cells];
	ranges[2 * q] = b;
	ranges[2 * q + 1] = class_start[(size_t)(q + 1) * slot_cells] - b;
}

// The policy table and the world record, passed by value (kernel arguments are copied at
// launch, so the call needs no host buffer that outlives it and no synchronisation).
__global__ void k_noalloc_put_policy(DevPolicy p, DevPolicy* dst) {
	if (threadIdx.x == 0 && blockIdx.x == 0) *dst = p;
}
struct NaSlots {
	int slot[NUM_CLS];
};
__global__ void k_noalloc_put_world(WorldInfo w, WorldInfo* dw, NaSlots s, int* dslot) {
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		*dw = w;
		for (int q = 0; q < NUM_CLS; ++q) dslot[q] = s.slot[q];
	}
}

int na_fail(char* err, int err_len, int code, const char* msg) {
	if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg);
	return code;
}

} // namespace

extern "C" int umans_step_noalloc(const umans_world_desc* world, const umans_policy* policies, int n_policies,
                                  const umans_agents* agents, int n, int nsteps, void* stream, char* err,
                                  int err_len) {
	if (err && err_len > 0) err[0] = '\0';
	if (!world || !policies || !agents) return na_fail(err, err_len, UMANS_EINVAL, "null argument");
	if (check_world(*world)) return na_fail(err, err_len, UMANS_EINVAL, "bad world description");
	if (world->max_neighbors < 1 || world->max_neighbors > kUncappedMax)
		return na_fail(err, err_len, UMANS_EINVAL, "max_neighbors must be 1..128");
	if (n_policies < 1 || n_policies > kNoAllocMaxPolicies)
		return na_fail(err, err_len, UMANS_EINVAL, "need 1..16 policies");
	if (n < 0 || n > UMANS_NOALLOC_MAX_AGENTS)
		return na_fail(err, err_len, UMANS_EINVAL, "agent count beyond UMANS_NOALLOC_MAX_AGENTS");
	if (nsteps < 0) return na_fail(err, err_len, UMANS_EINVAL, "negative step count");
	if (n > 0 && (!agents->id || !agents->pos || !agents->vel || !agents->goal || !agents->viewdir ||
	              !agents->radius || !agents->pref_speed || !agents->max_speed || !agents->max_accel ||
	              !agents->policy))
		return na_fail(err, err_len, UMANS_EINVAL, "missing agent array");
	if (agents->world) return na_fail(err, err_len, UMANS_EINVAL, "a single world: the world array must be null");
	const int kcap = world->max_neighbors;
	const int kmax = kmax_bucket(kcap);
	if (world->block_size > 0 && !block_supported(kmax, world->block_size))
		return na_fail(err, err_len, UMANS_EINVAL, "unsupported block size");
	// the policy table, with the classes it uses
	DevPolicy hp[kNoAllocMaxPolicies];
	float max_range = 0.0f;
	bool cls_used[NUM_CLS] = {};
	for (int i = 0; i < n_policies; ++i) {
		std::string e;
		if (make_dev_policy(policies[i], hp[i], e)) return na_fail(err, err_len, UMANS_EINVAL, e.c_str());
		max_range = std::max(max_range, hp[i].query_range);
		cls_used[hp[i].cls] = true;
	}
	int cls_slot[NUM_CLS] = {};
	int nslots = 0;
	for (int q = 0; q < NUM_CLS; ++q) cls_slot[q] = cls_used[q] ? nslots++ : 0;
	WorldInfo wi;
	const float cell = world->cell_size > 0.0f ? world->cell_size : std::max(max_range, 0.25f);
	const long long cells = world_info(*world, cell, n, wi);
	if (cells > kNoAllocMaxCells) return na_fail(err, err_len, UMANS_EINVAL, "grid beyond the fixed maximum");
	if (n == 0 || nsteps == 0) return UMANS_OK;

	// scratch's address on the current device, looked up once per device
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
		return na_fail(err, err_len, UMANS_EHIP, "no usable current device");
	unsigned char* base = nullptr;
	{
		std::lock_guard<std::mutex> lock(g_scratch_mu);
		if (!g_scratch_ptr[dev]) {
			void* p = nullptr;
			const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
			if (e != hipSuccess || !p) return na_fail(err, err_len, UMANS_EHIP, hipGetErrorString(e));
			g_scratch_ptr[dev] = p;
		}
		base = (unsigned char*)g_scratch_ptr[dev];
	}
	hipStream_t st = (hipStream_t)stream;
	DevPolicy* dpol = (DevPolicy*)(base + NaLayout::pol);
	WorldInfo* dworld = (WorldInfo*)(base + NaLayout::world);
	GridInfo* dgrid = (GridInfo*)(base + NaLayout::grid);
	int* dbbox = (int*)(base + NaLayout::bbox);
	unsigned long long* dstats = (unsigned long long*)(base + NaLayout::stats);
	int* dslot = (int*)(base + NaLayout::cls_slot);
	unsigned* dranges = (unsigned*)(base + NaLayout::ranges);
	unsigned* cell_count = (unsigned*)(base + NaLayout::cell_count);
	unsigned* cell_start = (unsigned*)(base + NaLayout::cell_start);
	unsigned* cell_first = (unsigned*)(base + NaLayout::cell_first);
	unsigned* scan = (unsigned*)(base + NaLayout::scan);
	unsigned* cell_of = (unsigned*)(base + NaLayout::cell_of);
	int* class_list = (int*)(base + NaLayout::class_list);
	const State X = na_state(base + NaLayout::state0);
	const State Y = na_state(base + NaLayout::state1);
	for (int i = 0; i < n_policies; ++i)
		hipLaunchKernelGGL(k_noalloc_put_policy, dim3(1), dim3(64), 0, st, hp[i], dpol + i);
	NaSlots slots;
	for (int q = 0; q < NUM_CLS; ++q) slots.slot[q] = cls_slot[q];
	hipLaunchKernelGGL(k_noalloc_put_world, dim3(1), dim3(64), 0, st, wi, dworld, slots, dslot);
	hipLaunchKernelGGL(k_noalloc_pack, dim3(ceil_div(n, 256)), dim3(256), 0, st, *agents, n, n_policies, X);

	const int total_cells = (int)cells;
	const int scan_n = nslots * total_cells + 1;
	const int nb = ceil_div(scan_n, kScanTile);
	const int bs = world->block_size > 0 ? world->block_size : default_block(kmax);
	constexpr int stat_words = 2 * kStatCopies * kStatLine;
	for (int s = 0; s < nsteps; ++s) {
		// the counters and the scan's ticket and statuses start at zero (every step: a
		// previous call may have stopped between the scan and the reorder)
		hipLaunchKernelGGL(k_fill, dim3(ceil_div(scan_n, 256)), dim3(256), 0, st, cell_count, scan_n, 0u);
		hipLaunchKernelGGL(k_fill, dim3(ceil_div(nb + 1, 256)), dim3(256), 0, st, scan, nb + 1, 0u);
		if (wi.world_type != 1) {
			hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, st, dbbox, 1);
			hipLaunchKernelGGL(k_bbox, dim3(ceil_div(n, 256)), dim3(256), 0, st, X.pos, X.meta, dworld, n, dbbox);
		}
		hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, st, dworld, dbbox, dgrid, 1, nullptr, 0, nullptr, 0, 0);
		hipLaunchKernelGGL(k_count, dim3(ceil_div(n, 256)), dim3(256), 0, st, X.pos, X.meta, n, dworld, dgrid, dpol,
		                   dslot, total_cells, cell_count, cell_of);
		hipLaunchKernelGGL(k_scan_lookback, dim3(nb), dim3(kScanThreads), 0, st, cell_count, cell_start, scan, scan + 1,
		                   scan_n, nb, (unsigned*)dstats, stat_words, nullptr, 0, kScanSpinLimit);
		const unsigned* cf = nslots > 1 ? cell_first : cell_start;
		const int nr = std::max(nslots > 1 ? std::max(n, total_cells + 1) : n, nb + 1);
		hipLaunchKernelGGL(k_reorder, dim3(ceil_div(nr, 256)), dim3(256), 0, st, X, Y, n, cell_of, cell_count, cell_start,
		                   nslots, total_cells, cell_first, class_list, scan, nb + 1, dworld, nullptr);
		hipLaunchKernelGGL(k_class_ranges, dim3(1), dim3(64), 0, st, cell_start, total_cells, nslots, dranges);
		StepArgs a{};
		a.in = Y;
		a.out = X;
		a.pol = dpol;
		a.worlds = dworld;
		a.grids = dgrid;
		a.next_grids = dgrid;
		a.cell_start = cf;
		a.class_list = class_list;
		a.n = n;
		a.kcap = kcap;
		a.stats = dstats;
		a.tier_final = 1;
		for (int qi = 0; qi < NUM_CLS; ++qi) {
			const int q = kLaunchOrder[qi];
			if (!cls_used[q]) continue;
			StepArgs aq = a;
			aq.dev_base = dranges + 2 * cls_slot[q];
			aq.agent_count = dranges + 2 * cls_slot[q] + 1;
			// lists longer than 32: one launch at full capacity (sorted insertion)
			launch_step(q, kmax, kmax <= 32 ? class_block(q, kmax, bs, nullptr) : 64, aq, st);
		}
	}
	hipLaunchKernelGGL(k_noalloc_unpack, dim3(ceil_div(n, 256)), dim3(256), 0, st, X, n, *agents);
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess) return na_fail(err, err_len, UMANS_EHIP, hipGetErrorString(e));
	return UMANS_OK;
}
