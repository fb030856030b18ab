// The reference's embedding API (src/Library/APIFunctions.cpp:35-205) over the C ABI of
// umans_gpu.h: one process-wide simulation, loaded from a scenario file, stepped on the
// GPU. Host code only; the population calls are queued by the library and applied on
// the device at the next step or read-back.
#include "umans_api.h"
#include "umans_gpu.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <unordered_map>
#include <vector>

namespace {

struct Simulation {
	umans_ctx* ctx = nullptr;
	float dt = 0.0f;
	std::unordered_map<int, int> policy_index;   // XML policy id -> packed policy
	umans_agent_record* records = nullptr;       // pinned, GetAgentPositions' buffer
	int record_cap = 0;

	~Simulation() {
		if (records) (void)hipHostFree(records);
		umans_destroy(ctx);
	}
};

Simulation* g_sim = nullptr;

static_assert(sizeof(AgentData) == sizeof(umans_agent_record), "AgentData layout");

int device_from_env() {
	const char* e = std::getenv("UMANS_DEVICE");
	return e ? std::atoi(e) : 0;
}

template <typename T> std::vector<T> pick(const std::vector<T>& v, const std::vector<int>& idx, int width) {
	std::vector<T> out(idx.size() * width);
	for (size_t i = 0; i < idx.size(); ++i)
		for (int j = 0; j < width; ++j) out[i * width + j] = v[(size_t)idx[i] * width + j];
	return out;
}

bool report(umans_ctx* ctx, int rc, const char* what) {
	if (rc == UMANS_OK) return true;
	std::fprintf(stderr, "Error: %s failed: %s\n", what, ctx ? umans_last_error(ctx) : "");
	return false;
}

// CrowdSimulator::FromConfigFile as the reference's StartSimulation uses it: policies,
// agents in file order (start time 0 now, later ones scheduled), remove_at_goal flags.
Simulation* load(const char* path) {
	umans_scenario* sc = nullptr;
	char err[1024] = {0};
	if (umans_scenario_load(path, &sc, err, (int)sizeof(err)) != UMANS_OK) {
		std::fprintf(stderr, "%s\n", err);
		return nullptr;
	}
	umans_scenario_info info{};
	umans_scenario_get_info(sc, &info);
	const int n = info.n_agents, np = info.n_policies;
	std::vector<umans_policy> pol(np);
	std::vector<int32_t> xml_ids(np), id(n), policy(n), rm(n);
	std::vector<float> pos(2 * n), vel(2 * n), goal(2 * n), view(2 * n), rad(n), ps(n), ms(n), ma(n), st(n);
	std::vector<uint32_t> rng(n);
	umans_agents a{};
	a.id = id.data(); a.pos = pos.data(); a.vel = vel.data(); a.goal = goal.data(); a.viewdir = view.data();
	a.radius = rad.data(); a.pref_speed = ps.data(); a.max_speed = ms.data(); a.max_accel = ma.data();
	a.policy = policy.data(); a.rng = rng.data();
	const int grc = umans_scenario_get(sc, pol.data(), xml_ids.data(), &a, st.data(), rm.data());
	// obstacle polygons: vertex counts, then the vertices
	std::vector<int32_t> ocount(info.n_obstacles);
	int nverts = 0;
	if (info.n_obstacles > 0) {
		umans_scenario_obstacles(sc, ocount.data(), nullptr);
		for (int32_t v : ocount) nverts += v;
	}
	std::vector<float> opts(2 * (size_t)nverts);
	if (info.n_obstacles > 0) umans_scenario_obstacles(sc, ocount.data(), opts.data());
	umans_scenario_free(sc);
	if (grc != UMANS_OK) {
		std::fprintf(stderr, "Error: could not read the scenario %s\n", path);
		return nullptr;
	}
	Simulation* s = new Simulation();
	s->dt = info.delta_time;
	for (int i = 0; i < np; ++i) s->policy_index[xml_ids[i]] = i;
	umans_world_desc d{};
	d.world_type = info.world_type;
	d.width = info.width;
	d.height = info.height;
	d.delta_time = info.delta_time;
	d.device = device_from_env();
	if (!report(nullptr, umans_create(&d, &s->ctx), "umans_create") ||
	    !report(s->ctx, umans_set_policies(s->ctx, pol.data(), np), "loading the policies")) {
		delete s;
		return nullptr;
	}
	if (info.n_obstacles > 0 &&
	    !report(s->ctx, umans_set_obstacles(s->ctx, info.n_obstacles, ocount.data(), opts.data()), "loading the obstacles")) {
		delete s;
		return nullptr;
	}
	std::vector<int> now, later;
	for (int i = 0; i < n; ++i) (st[i] <= 0.0f ? now : later).push_back(i);
	auto subset = [&](const std::vector<int>& idx, std::vector<std::vector<float>>& f, std::vector<int32_t>& iv,
	                  std::vector<int32_t>& pv, std::vector<uint32_t>& rv) {
		f = {pick(pos, idx, 2), pick(vel, idx, 2), pick(goal, idx, 2), pick(view, idx, 2), pick(rad, idx, 1),
		     pick(ps, idx, 1), pick(ms, idx, 1), pick(ma, idx, 1)};
		iv = pick(id, idx, 1);
		pv = pick(policy, idx, 1);
		rv = pick(rng, idx, 1);
		umans_agents b{};
		b.id = iv.data(); b.pos = f[0].data(); b.vel = f[1].data(); b.goal = f[2].data(); b.viewdir = f[3].data();
		b.radius = f[4].data(); b.pref_speed = f[5].data(); b.max_speed = f[6].data(); b.max_accel = f[7].data();
		b.policy = pv.data(); b.rng = rv.data();
		return b;
	};
	std::vector<std::vector<float>> f0, f1;
	std::vector<int32_t> i0, i1, p0, p1;
	std::vector<uint32_t> r0, r1;
	umans_agents b0 = subset(now, f0, i0, p0, r0);
	if (!report(s->ctx, umans_upload(s->ctx, &b0, (int)now.size()), "uploading the agents")) {
		delete s;
		return nullptr;
	}
	std::vector<uint8_t> rm_now;
	for (int i : now) rm_now.push_back(rm[i] ? 1 : 0);
	if (!report(s->ctx, umans_set_remove_at_goal(s->ctx, (int)now.size(), i0.data(), rm_now.data()), "remove_at_goal")) {
		delete s;
		return nullptr;
	}
	if (!later.empty()) {
		umans_agents b1 = subset(later, f1, i1, p1, r1);
		std::vector<float> st1 = pick(st, later, 1);
		std::vector<uint8_t> rm1;
		for (int i : later) rm1.push_back(rm[i] ? 1 : 0);
		std::vector<int32_t> out((size_t)later.size());
		if (!report(s->ctx, umans_add_agents(s->ctx, &b1, (int)later.size(), st1.data(), rm1.data(),
		                                     UMANS_ADD_KEEP_IDS, out.data()),
		            "scheduling the agents")) {
			delete s;
			return nullptr;
		}
	}
	return s;
}

} // namespace

extern "C" {

bool StartSimulation(const char* configFileName, int nrThreads) {
	(void)nrThreads;
	delete g_sim;
	g_sim = configFileName ? load(configFileName) : nullptr;
	return g_sim != nullptr;
}

bool GetSimulationTimeStep(float& result_dt) {
	if (!g_sim) return false;
	result_dt = g_sim->dt;
	return true;
}

bool DoSimulationSteps(int nrSteps) {
	if (!g_sim) return false;
	return report(g_sim->ctx, umans_step(g_sim->ctx, nrSteps), "DoSimulationSteps");
}

bool GetAgentPositions(AgentData*& result_agentData, int& result_nrAgents) {
	if (!g_sim) return false;
	Simulation& s = *g_sim;
	const int n = umans_num_agents(s.ctx);   // the list length, pending changes included
	if (n > s.record_cap) {
		if (s.records) (void)hipHostFree(s.records);
		s.records = nullptr;
		s.record_cap = 0;
		int cap = 1;
		while (cap < n) cap <<= 1;
		if (hipHostMalloc((void**)&s.records, (size_t)cap * sizeof(umans_agent_record)) != hipSuccess) {
			s.records = nullptr;
			std::fprintf(stderr, "Error: GetAgentPositions could not allocate %d records\n", cap);
			return false;
		}
		s.record_cap = cap;
	}
	if (!report(s.ctx, umans_download_records(s.ctx, s.records, s.record_cap), "GetAgentPositions")) return false;
	result_agentData = reinterpret_cast<AgentData*>(s.records);
	result_nrAgents = n;
	return true;
}

bool SetAgentPositions(AgentData* agentData, int nrAgents) {
	if (!g_sim) return false;
	if (nrAgents <= 0) return true;
	std::vector<int32_t> ids(nrAgents);
	std::vector<float> pos(2 * nrAgents), vel(2 * nrAgents), view(2 * nrAgents);
	for (int i = 0; i < nrAgents; ++i) {
		const AgentData& d = agentData[i];
		ids[i] = d.id;
		pos[2 * i] = d.position_x; pos[2 * i + 1] = d.position_y;
		vel[2 * i] = d.velocity_x; vel[2 * i + 1] = d.velocity_y;
		view[2 * i] = d.viewingDirection_x; view[2 * i + 1] = d.viewingDirection_y;
	}
	return report(g_sim->ctx, umans_set_agent_states(g_sim->ctx, nrAgents, ids.data(), pos.data(), vel.data(),
	                                                  view.data()), "SetAgentPositions");
}

bool AddAgent(float x, float y, float radius, float prefSpeed, float maxSpeed, float maxAcceleration, int policyID,
              int& result_id, int desiredID) {
	if (!g_sim) return false;
	const auto it = g_sim->policy_index.find(policyID);
	if (it == g_sim->policy_index.end()) return false;
	int32_t id = desiredID >= 0 ? desiredID : -1;
	int32_t pol = it->second;
	float pos[2] = {x, y};
	umans_agents a{};
	a.id = &id; a.pos = pos; a.radius = &radius; a.pref_speed = &prefSpeed; a.max_speed = &maxSpeed;
	a.max_accel = &maxAcceleration; a.policy = &pol;
	int32_t out = -1;
	if (!report(g_sim->ctx, umans_add_agents(g_sim->ctx, &a, 1, nullptr, nullptr, 0, &out), "AddAgent")) return false;
	result_id = out;
	return true;
}

bool RemoveAgent(int id) {
	if (!g_sim) return false;
	return umans_remove_agent(g_sim->ctx, id) == UMANS_OK;
}

bool SetAgentGoal(int id, float x, float y) {
	if (!g_sim) return false;
	return umans_set_agent_goal(g_sim->ctx, id, x, y) == UMANS_OK;
}

bool CleanUp() {
	if (!g_sim) return false;
	delete g_sim;
	g_sim = nullptr;
	return true;
}

} // extern "C"
