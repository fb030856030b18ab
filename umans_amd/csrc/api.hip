// libumans_gpu: context management and the C ABI of include/umans_gpu.h.
// The kernels are in kernels.h; the registry in registry.cpp.
#include "umans_gpu.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace umans;

namespace {

constexpr int kMaxPolicies = 1024;
constexpr int kUncappedMax = 128;

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
		hipFree(s.pos); hipFree(s.vel); hipFree(s.goal); hipFree(s.viewdir);
		hipFree(s.pref); hipFree(s.prm); hipFree(s.meta);
		s = State{};
	}
};

} // namespace

struct umans_ctx {
	umans_world_desc desc{};
	int device = 0;
	hipStream_t stream = nullptr;
	std::string err;

	std::vector<DevPolicy> hpol;
	DevPolicy* dpol = nullptr;
	float max_range = 0.0f;

	int n = 0, cap = 0;
	DeviceState X, Y;            // X: state between steps, Y: cell-sorted state of a step
	unsigned* cell_count = nullptr;
	unsigned* cell_start = nullptr;
	unsigned* scan_sums = nullptr;
	unsigned* cell_of = nullptr;
	unsigned* rank = nullptr;
	int cell_cap = 0;
	int scan_cells = 0;          // cells the scan covers (toric: the grid; infinite: cell_cap)
	GridInfo* dgrid = nullptr;
	int* dbbox = nullptr;
	unsigned long long* dstats = nullptr;
	int kcap = 0;

	int record = 0;
	int* rec_count = nullptr;
	int* rec_id = nullptr;
	double* rec_d2 = nullptr;
	int* rec_seg = nullptr;
	int rec_stride = 0;
	bool have_record = false;

	long long steps = 0;
	double sim_time = 0.0;

	// per-phase timers: one event triple per profiled step, read back lazily
	int profiling = 0;
	std::vector<hipEvent_t> ev;          // 3 per pending step
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
	void free_agents() {
		X.release(); Y.release();
		hipFree(cell_count); hipFree(cell_start); hipFree(scan_sums); hipFree(cell_of); hipFree(rank);
		hipFree(rec_count); hipFree(rec_id); hipFree(rec_d2); hipFree(rec_seg);
		cell_count = cell_start = scan_sums = cell_of = rank = nullptr;
		rec_count = rec_id = rec_seg = nullptr;
		rec_d2 = nullptr;
		n = cap = cell_cap = 0;
	}
	// read back and accumulate the pending event triples
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
			hipEventCreate(&e);
			ev.push_back(e);
		}
		return ev[ev_used++];
	}
	void free_events() {
		for (hipEvent_t e : ev) hipEventDestroy(e);
		ev.clear();
		ev_used = 0;
	}
};

namespace {

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
	if (p.method == UMANS_METHOD_GRADIENT) {
		d.mode = MODE_GRADIENT;
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
	if (d.mode == MODE_SAMPLE) {
		if (s.s_type == 0 && (s.s_speed_samples < 1 || s.s_angle_samples < 1)) {
			err = "regular sampling needs SpeedSamples >= 1 and AngleSamples >= 1";
			return UMANS_EINVAL;
		}
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

int kmax_bucket(int kcap) {
	if (kcap <= 16) return 16;
	if (kcap <= 32) return 32;
	if (kcap <= 64) return 64;
	return 128;
}

template <int KMAX, int BS>
void launch_step_t(const StepArgs& a, hipStream_t st) {
	const int blocks = (a.n + BS - 1) / BS;
	if (blocks > 0) hipLaunchKernelGGL((k_step<KMAX, BS>), dim3(blocks), dim3(BS), 0, st, a);
}

void launch_step(int kmax, const StepArgs& a, hipStream_t st) {
	switch (kmax) {
	case 16: launch_step_t<16, 128>(a, st); break;
	case 32: launch_step_t<32, 128>(a, st); break;
	case 64: launch_step_t<64, 64>(a, st); break;
	default: launch_step_t<128, 64>(a, st); break;
	}
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Toric grid: fixed for the world; cells at least `cs` wide.
GridInfo toric_grid(float W, float H, float cs, int cap) {
	GridInfo g{};
	long long nx = std::max(1LL, (long long)std::floor((double)W / cs));
	long long ny = std::max(1LL, (long long)std::floor((double)H / cs));
	while (nx * ny > cap) {
		if (nx >= ny) nx = std::max(1LL, nx / 2); else ny = std::max(1LL, ny / 2);
	}
	g.ox = -0.5f * W; g.oy = -0.5f * H;
	g.ncx = (int)nx; g.ncy = (int)ny;
	g.invx = (float)((double)nx / W); g.invy = (float)((double)ny / H);
	g.ncells = g.ncx * g.ncy;
	return g;
}

int enqueue_step(umans_ctx* c) {
	const int n = c->n;
	hipStream_t st = c->stream;
	const int cells = c->scan_cells;  // counters cleared and scanned: cells + 1
	const float cs = c->desc.cell_size > 0.0f ? c->desc.cell_size : std::max(c->max_range, 0.5f);

	hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
	if (c->profiling) {
		if (c->ev_used + 3 > 3 * 256) {      // bound the pending events
			const int rc = c->drain_profile();
			if (rc) return rc;
		}
		e0 = c->next_event(); e1 = c->next_event(); e2 = c->next_event();
		hipEventRecord(e0, st);
	}
	hipLaunchKernelGGL(k_fill, dim3(ceil_div(cells + 1, 256)), dim3(256), 0, st, c->cell_count, cells + 1, 0u);
	hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (unsigned*)c->dstats, 4, 0u);
	if (c->desc.world_type != 1) {
		hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, st, c->dbbox);
		hipLaunchKernelGGL(k_bbox, dim3(std::min(ceil_div(n, 256), 1024)), dim3(256), 0, st, c->X.s.pos, n, c->dbbox);
		hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, st, c->dbbox, cs, c->cell_cap, c->dgrid);
	}
	hipLaunchKernelGGL(k_count, dim3(ceil_div(n, 256)), dim3(256), 0, st, c->X.s.pos, n, c->dgrid,
	                   c->cell_count, c->cell_of, c->rank);
	const int scan_n = cells + 1;
	const int nb = ceil_div(scan_n, kScanTile);
	hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanThreads), 0, st, c->cell_count, c->cell_start, c->scan_sums, scan_n);
	hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, c->scan_sums, nb);
	hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, st, c->cell_start, c->scan_sums, scan_n);
	hipLaunchKernelGGL(k_reorder, dim3(ceil_div(n, 256)), dim3(256), 0, st, c->X.s, c->Y.s, n, c->cell_of, c->rank,
	                   c->cell_start);

	StepArgs a{};
	a.in = c->Y.s;
	a.out = c->X.s;
	a.pol = c->dpol;
	a.grid = c->dgrid;
	a.cell_start = c->cell_start;
	a.n = n;
	a.world_type = c->desc.world_type;
	a.W = c->desc.width;
	a.H = c->desc.height;
	a.dt = c->desc.delta_time;
	a.kcap = c->kcap;
	if (c->record) {
		a.rec_count = c->rec_count; a.rec_id = c->rec_id; a.rec_d2 = c->rec_d2; a.rec_seg = c->rec_seg;
		a.rec_stride = c->rec_stride;
	}
	a.stats = c->dstats;
	if (e1) hipEventRecord(e1, st);
	launch_step(kmax_bucket(c->kcap), a, st);
	if (e2) hipEventRecord(e2, st);
	return c->hip(hipGetLastError(), "step launch");
}

} // namespace

extern "C" {

int umans_create(const umans_world_desc* desc, umans_ctx** out) {
	if (!desc || !out) return UMANS_EINVAL;
	*out = nullptr;
	if (desc->world_type != 0 && desc->world_type != 1) return UMANS_EINVAL;
	if (!(desc->delta_time > 0.0f)) return UMANS_EINVAL;   // crowdSimulator.cpp:575-584
	if (desc->world_type == 1 && !(desc->width > 0.0f && desc->height > 0.0f)) return UMANS_EINVAL;
	umans_ctx* c = new umans_ctx();
	c->desc = *desc;
	c->device = desc->device;
	c->kcap = desc->max_neighbors > 0 ? desc->max_neighbors : kUncappedMax;
	if (c->kcap > kUncappedMax) {
		delete c;
		return UMANS_EINVAL;
	}
	if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
	    dalloc(&c->dgrid, 1) != hipSuccess || dalloc(&c->dbbox, 4) != hipSuccess ||
	    dalloc(&c->dstats, 2) != hipSuccess || dalloc(&c->dpol, kMaxPolicies) != hipSuccess) {
		umans_destroy(c);
		return UMANS_EHIP;
	}
	*out = c;
	return UMANS_OK;
}

void umans_destroy(umans_ctx* c) {
	if (!c) return;
	hipSetDevice(c->device);
	if (c->stream) hipStreamSynchronize(c->stream);
	c->free_agents();
	hipFree(c->dgrid); hipFree(c->dbbox); hipFree(c->dstats); hipFree(c->dpol);
	c->free_events();
	if (c->stream) hipStreamDestroy(c->stream);
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
	hipSetDevice(c->device);
	const int rc = c->hip(hipMemcpyAsync(c->dpol, v.data(), n * sizeof(DevPolicy), hipMemcpyHostToDevice, c->stream),
	                      "policy upload");
	if (rc) return rc;
	c->hpol.swap(v);
	c->max_range = mr;
	return c->hip(hipStreamSynchronize(c->stream), "policy upload");
}

int umans_upload(umans_ctx* c, const umans_agents* a, int n) {
	if (!c) return UMANS_EINVAL;
	if (c->hpol.empty()) return c->fail(UMANS_ESTATE, "set policies before uploading agents");
	if (!a || n < 0 || n > (int)kIndexMask) return c->fail(UMANS_EINVAL, "bad agent count");
	if (n > 0 && (!a->id || !a->pos || !a->goal || !a->radius || !a->pref_speed || !a->max_speed || !a->max_accel || !a->policy))
		return c->fail(UMANS_EINVAL, "missing agent array");
	const int npol = (int)c->hpol.size();
	std::vector<float2> pos(n), vel(n), goal(n), view(n);
	std::vector<float4> prm(n);
	std::vector<int4> meta(n);
	for (int i = 0; i < n; ++i) {
		if (a->policy[i] < 0 || a->policy[i] >= npol)
			return c->fail(UMANS_EINVAL, "agent " + std::to_string(i) + " has an unknown policy id");
		pos[i] = make_float2(a->pos[2 * i], a->pos[2 * i + 1]);
		vel[i] = a->vel ? make_float2(a->vel[2 * i], a->vel[2 * i + 1]) : make_float2(0.0f, 0.0f);
		goal[i] = make_float2(a->goal[2 * i], a->goal[2 * i + 1]);
		if (a->viewdir) view[i] = make_float2(a->viewdir[2 * i], a->viewdir[2 * i + 1]);
		else {
			// Agent::setGoal looks at the goal (agent.cpp:120-127)
			float dx = goal[i].x - pos[i].x, dy = goal[i].y - pos[i].y;
			const float m = std::sqrt(dx * dx + dy * dy);
			view[i] = m > 0.0f ? make_float2(dx / m, dy / m) : make_float2(0.0f, 0.0f);
		}
		prm[i] = make_float4(a->radius[i], a->pref_speed[i], a->max_speed[i], a->max_accel[i]);
		uint32_t r;
		if (a->rng) r = a->rng[i];
		else {
			// std::default_random_engine(id) = minstd_rand0 seeded with the ID (agent.cpp:38)
			r = (uint32_t)((uint32_t)a->id[i] % 2147483647u);
			if (r == 0) r = 1;
		}
		meta[i] = make_int4(a->id[i], a->policy[i], i, (int)r);
	}
	hipSetDevice(c->device);
	if (n > c->cap) {
		c->free_agents();
		const int cap = n;
		int rc;
		if ((rc = c->hip(c->X.alloc(cap), "alloc state"))) return rc;
		if ((rc = c->hip(c->Y.alloc(cap), "alloc state"))) return rc;
		const int cells = std::max(4 * cap, 4096);
		if ((rc = c->hip(dalloc(&c->cell_count, cells + 1), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->cell_start, cells + 1), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->scan_sums, ceil_div(cells + 1, kScanTile) + 1), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->cell_of, cap), "alloc grid"))) return rc;
		if ((rc = c->hip(dalloc(&c->rank, cap), "alloc grid"))) return rc;
		c->cap = cap;
		c->cell_cap = cells;
	}
	c->n = n;
	c->have_record = false;
	if (c->desc.world_type == 1) {
		const float cs = c->desc.cell_size > 0.0f ? c->desc.cell_size : std::max(c->max_range, 0.5f);
		const GridInfo g = toric_grid(c->desc.width, c->desc.height, cs, c->cell_cap);
		int rc = c->hip(hipMemcpyAsync(c->dgrid, &g, sizeof(g), hipMemcpyHostToDevice, c->stream), "grid upload");
		if (rc) return rc;
		c->scan_cells = g.ncells;
	} else {
		c->scan_cells = c->cell_cap;
	}
	const State& s = c->X.s;
	int rc;
	if ((rc = c->hip(hipMemcpyAsync(s.pos, pos.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.vel, vel.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.goal, goal.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.viewdir, view.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.prm, prm.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	if ((rc = c->hip(hipMemcpyAsync(s.meta, meta.data(), n * sizeof(int4), hipMemcpyHostToDevice, c->stream), "upload"))) return rc;
	return c->hip(hipStreamSynchronize(c->stream), "upload");
}

int umans_num_agents(const umans_ctx* c) { return c ? c->n : 0; }

int umans_set_record_neighbors(umans_ctx* c, int enable) {
	if (!c) return UMANS_EINVAL;
	if (enable && c->cap > 0 && !c->rec_count) {
		hipSetDevice(c->device);
		const int stride = kmax_bucket(c->kcap);
		int rc;
		if ((rc = c->hip(dalloc(&c->rec_count, c->cap), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_id, (size_t)c->cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_d2, (size_t)c->cap * stride), "alloc record"))) return rc;
		if ((rc = c->hip(dalloc(&c->rec_seg, (size_t)c->cap * stride), "alloc record"))) return rc;
		c->rec_stride = stride;
	}
	if (enable && !c->rec_count) return c->fail(UMANS_ESTATE, "upload agents before enabling the neighbour record");
	c->record = enable ? 1 : 0;
	return UMANS_OK;
}

int umans_step(umans_ctx* c, int nsteps) {
	if (!c) return UMANS_EINVAL;
	if (nsteps < 0) return c->fail(UMANS_EINVAL, "negative step count");
	if (c->hpol.empty()) return c->fail(UMANS_ESTATE, "no policies");
	hipSetDevice(c->device);
	for (int s = 0; s < nsteps; ++s) {
		if (c->n > 0) {
			const int rc = enqueue_step(c);
			if (rc) return rc;
		}
		c->steps += 1;
		c->sim_time += (double)c->desc.delta_time;   // WorldBase::time_ (worldBase.cpp:122)
	}
	if (c->record && nsteps > 0) c->have_record = true;
	return UMANS_OK;
}

int umans_set_profiling(umans_ctx* c, int enable) {
	if (!c) return UMANS_EINVAL;
	hipSetDevice(c->device);
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
	hipSetDevice(c->device);
	const int rc = c->drain_profile();
	if (rc) return rc;
	out->steps = c->prof_steps;
	out->grid_ms = c->prof_grid_ms;
	out->step_ms = c->prof_step_ms;
	return UMANS_OK;
}

int umans_synchronize(umans_ctx* c) {
	if (!c) return UMANS_EINVAL;
	hipSetDevice(c->device);
	return c->hip(hipStreamSynchronize(c->stream), "synchronize");
}

int umans_get_stats(umans_ctx* c, umans_stats* out) {
	if (!c || !out) return UMANS_EINVAL;
	hipSetDevice(c->device);
	unsigned long long h[2] = {0, 0};
	int rc = c->hip(hipMemcpyAsync(h, c->dstats, sizeof(h), hipMemcpyDeviceToHost, c->stream), "stats");
	if (rc) return rc;
	if ((rc = c->hip(hipStreamSynchronize(c->stream), "stats"))) return rc;
	out->steps = c->steps;
	out->overflow_agents = (int64_t)h[0];
	out->max_list = (int64_t)h[1];
	out->sim_time = c->sim_time;
	return UMANS_OK;
}

int umans_download(umans_ctx* c, umans_agents* a) {
	if (!c || !a) return UMANS_EINVAL;
	hipSetDevice(c->device);
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
		if (a->policy) a->policy[i] = meta[k].y;
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
	return UMANS_OK;
}

int umans_dump_neighbors(umans_ctx* c, int stride, int32_t* count, int32_t* ids, double* d2, int32_t* seg) {
	if (!c || !count || stride < 0 || (stride > 0 && (!ids || !d2 || !seg))) return UMANS_EINVAL;
	if (!c->have_record) return c->fail(UMANS_ESTATE, "no neighbour record: enable it and run a step");
	hipSetDevice(c->device);
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

} // extern "C"
