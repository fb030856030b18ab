// Test-only driver around the UMANS reference engine, compiled together with the
// reference's own sources by oracle/Makefile into oracle/_ref/libumans_ref.so.
//
// This file is test infrastructure: only tests/, the smoke check and bench.py's
// CPU-baseline leg load the resulting library. It exposes a plain C ABI so that
// Python (ctypes) can
//   * load a scenario XML through CrowdSimulator::FromConfigFile
//     (src/Engine/core/crowdSimulator.cpp:516-647),
//   * add agents programmatically through WorldBase::AddAgent
//     (src/Engine/core/worldBase.cpp:161-187) + Agent::setGoal (agent.cpp:120-127),
//   * run WorldBase::DoStep (worldBase.cpp:79-128) and time it,
//   * read back per-agent state and the neighbour lists each agent computed in
//     the last step (Agent::getNeighbors, agent.h:155; PhantomAgent, worldBase.h:40-79).
// Nothing here changes the reference's arithmetic.

#include <core/crowdSimulator.h>
#include <core/worldBase.h>
#include <core/agent.h>
#include <core/policy.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

void* ref_load(const char* xml_path)
{
	return CrowdSimulator::FromConfigFile(xml_path);
}

void ref_destroy(void* sim)
{
	delete static_cast<CrowdSimulator*>(sim);
}

void ref_set_threads(void* sim, int n)
{
	static_cast<CrowdSimulator*>(sim)->GetWorld()->SetNumberOfThreads(n);
}

float ref_delta_time(void* sim)
{
	return static_cast<CrowdSimulator*>(sim)->GetWorld()->GetDeltaTime();
}

int ref_num_agents(void* sim)
{
	return (int)static_cast<CrowdSimulator*>(sim)->GetWorld()->GetNumberOfAgents();
}

// Adds n agents; returns the number that were added (0 if a policy id is unknown).
int ref_add_agents(void* sim, int n, const int* ids, const float* pos, const float* goal,
	const float* radius, const float* pref_speed, const float* max_speed, const float* max_accel,
	const int* policy_id)
{
	auto* cs = static_cast<CrowdSimulator*>(sim);
	for (int i = 0; i < n; ++i)
	{
		Agent::Settings s;
		s.radius_ = radius[i];
		s.preferred_speed_ = pref_speed[i];
		s.max_speed_ = max_speed[i];
		s.max_acceleration_ = max_accel[i];
		s.policy_ = cs->GetPolicy(policy_id[i]);
		if (s.policy_ == nullptr)
			return i;
		Agent* a = cs->GetWorld()->AddAgent(Vector2D(pos[2 * i], pos[2 * i + 1]), s,
			ids ? (size_t)ids[i] : std::numeric_limits<size_t>::max(), 0.0f);
		a->setGoal(Vector2D(goal[2 * i], goal[2 * i + 1]));
	}
	return n;
}

// Overrides velocities (and viewing directions) via Agent::setVelocity_ExternalApplication.
void ref_set_velocities(void* sim, int n, const float* vel, const float* viewdir)
{
	auto& agents = static_cast<CrowdSimulator*>(sim)->GetWorld()->GetAgents();
	for (int i = 0; i < n && i < (int)agents.size(); ++i)
		agents[i]->setVelocity_ExternalApplication(Vector2D(vel[2 * i], vel[2 * i + 1]),
			Vector2D(viewdir[2 * i], viewdir[2 * i + 1]));
}

// CrowdSimulator::StartCSVOutput (crowdSimulator.cpp:37-51); flushed when destroyed.
void ref_start_csv(void* sim, const char* dir, int flush_immediately)
{
	static_cast<CrowdSimulator*>(sim)->StartCSVOutput(dir, flush_immediately != 0);
}

void ref_step(void* sim, int nsteps)
{
	static_cast<CrowdSimulator*>(sim)->RunSimulationSteps(nsteps);
}

// Wall-clock seconds for nsteps calls of WorldBase::DoStep.
double ref_time_steps(void* sim, int nsteps)
{
	auto t0 = std::chrono::steady_clock::now();
	static_cast<CrowdSimulator*>(sim)->RunSimulationSteps(nsteps);
	auto t1 = std::chrono::steady_clock::now();
	return std::chrono::duration<double>(t1 - t0).count();
}

// Per-agent state in world list order. Any output pointer may be null.
void ref_get_state(void* sim, int* ids, float* pos, float* vel, float* viewdir, float* pref_vel,
	float* goal)
{
	const auto& agents = static_cast<CrowdSimulator*>(sim)->GetWorld()->GetAgents();
	for (size_t i = 0; i < agents.size(); ++i)
	{
		const Agent* a = agents[i];
		if (ids) ids[i] = (int)a->getID();
		if (pos) { pos[2 * i] = a->getPosition().x; pos[2 * i + 1] = a->getPosition().y; }
		if (vel) { vel[2 * i] = a->getVelocity().x; vel[2 * i + 1] = a->getVelocity().y; }
		if (viewdir) { viewdir[2 * i] = a->getViewingDirection().x; viewdir[2 * i + 1] = a->getViewingDirection().y; }
		if (pref_vel) { pref_vel[2 * i] = a->getPreferredVelocity().x; pref_vel[2 * i + 1] = a->getPreferredVelocity().y; }
		if (goal) { goal[2 * i] = a->getGoal().x; goal[2 * i + 1] = a->getGoal().y; }
	}
}

// Per-agent settings in world list order: radius, preferred / maximum speed, maximum
// acceleration, and the XML id of the agent's policy (searched in 0..max_policy_id).
void ref_get_settings(void* sim, float* radius, float* pref_speed, float* max_speed, float* max_accel,
	int* policy_xml_id, int max_policy_id)
{
	auto* cs = static_cast<CrowdSimulator*>(sim);
	const auto& agents = cs->GetWorld()->GetAgents();
	for (size_t i = 0; i < agents.size(); ++i)
	{
		const Agent* a = agents[i];
		radius[i] = a->getRadius();
		pref_speed[i] = a->getPreferredSpeed();
		max_speed[i] = a->getMaximumSpeed();
		max_accel[i] = a->getMaximumAcceleration();
		policy_xml_id[i] = -1;
		for (int p = 0; p <= max_policy_id; ++p)
			if (cs->GetPolicy(p) == a->getPolicy()) { policy_xml_id[i] = p; break; }
	}
}

// Total number of neighbour entries over all agents from the last step.
long long ref_total_neighbors(void* sim)
{
	const auto& agents = static_cast<CrowdSimulator*>(sim)->GetWorld()->GetAgents();
	long long t = 0;
	for (const Agent* a : agents)
		t += (long long)a->getNeighbors().first.size();
	return t;
}

// Flattened neighbour lists from the last step (list order as the reference
// stores it): counts[n], then per entry the neighbour ID, the phantom position
// and the phantom's float squared distance to the querying agent.
void ref_get_neighbors(void* sim, int* counts, int* nb_ids, float* nb_pos, float* nb_dsq)
{
	const auto& agents = static_cast<CrowdSimulator*>(sim)->GetWorld()->GetAgents();
	long long o = 0;
	for (size_t i = 0; i < agents.size(); ++i)
	{
		const auto& nl = agents[i]->getNeighbors().first;
		counts[i] = (int)nl.size();
		for (const auto& ph : nl)
		{
			nb_ids[o] = (int)ph.realAgent->getID();
			nb_pos[2 * o] = ph.GetPosition().x;
			nb_pos[2 * o + 1] = ph.GetPosition().y;
			nb_dsq[o] = ph.GetDistanceSquared();
			++o;
		}
	}
}

} // extern "C"
