/* The reference's embedding API, GPU-backed.
 *
 * Same symbols, signatures and return conventions as the reference's
 * src/Library/APIFunctions.h:30-113 (a C++ header: the functions are extern "C" but use
 * references and a default argument, as the reference's do). One simulation per
 * process, created by StartSimulation and ended by CleanUp; every function returns
 * false when no simulation is running. Exported by umans_amd/libumans_gpu.so, so an
 * application linked against the reference's libUMANS-Library.so can load this library
 * instead.
 *
 * Differences, each a limit of the GPU path:
 *   - nrThreads is ignored (no host threads take part in a step);
 *   - scenarios with cost functions the GPU path does not implement are refused
 *     (StartSimulation returns false and prints why); an agent keeps at most 32
 *     obstacle edges in range;
 *   - AddAgent returns false for a policy ID the scenario does not define (the
 *     reference would add an agent without a policy);
 *   - the GPU is device 0, or the one named by the environment variable UMANS_DEVICE.
 */
#ifndef UMANS_API_H
#define UMANS_API_H

extern "C" {

/* APIFunctions.h:33-49: the status of one agent, 28 bytes. */
struct AgentData {
	int id;
	float position_x;
	float position_y;
	float velocity_x;
	float velocity_y;
	float viewingDirection_x;
	float viewingDirection_y;
};

/* APIFunctions.h:54 (APIFunctions.cpp:47-66): load a scenario file. */
bool StartSimulation(const char* configFileName, int nrThreads);

/* APIFunctions.h:59 (APIFunctions.cpp:68-75) */
bool GetSimulationTimeStep(float& result_dt);

/* APIFunctions.h:65 (APIFunctions.cpp:77-84) */
bool DoSimulationSteps(int nrSteps);

/* APIFunctions.h:73 (APIFunctions.cpp:86-115): result_agentData points into a buffer
 * owned by the library (pinned host memory), valid until the next call of this
 * function or CleanUp; agents in the list order of WorldBase::GetAgents. */
bool GetAgentPositions(AgentData*& result_agentData, int& result_nrAgents);

/* APIFunctions.h:79 (APIFunctions.cpp:117-137): position, velocity and viewing
 * direction by ID; unknown IDs are skipped. */
bool SetAgentPositions(AgentData* agentData, int nrAgents);

/* APIFunctions.h:94 (APIFunctions.cpp:139-169): the goal is the start position. */
bool AddAgent(float x, float y, float radius, float prefSpeed, float maxSpeed, float maxAcceleration,
              int policyID, int& result_id, int desiredID = -1);

/* APIFunctions.h:100 (APIFunctions.cpp:171-178) */
bool RemoveAgent(int id);

/* APIFunctions.h:108 (APIFunctions.cpp:180-192) */
bool SetAgentGoal(int id, float x, float y);

/* APIFunctions.h:112 (APIFunctions.cpp:194-205) */
bool CleanUp();

}

#endif
