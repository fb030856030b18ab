// Cost-function registry (host only): the reference's registry names, XML parameter
// names, parameter read order and defaults, mapped to the packed umans_policy layout.
// CostFunctionFactory::RegisterAllCostFunctions (costFunctionFactory.cpp:41-61) and
// each cost function's parseParameters; defaults from the class headers.
#include "umans_gpu.h"

#include <cstring>

namespace {

struct Param { const char* name; int slot; float dflt; bool has_default; };
struct CostDef { const char* name; int kind; float range; int nparam; Param params[6]; int nslots; };

// Read order matters for SocialForcesAvoidance: "U0" is read into V0 after "V0"
// (SocialForcesAvoidance.cpp:142-150).
const CostDef kCosts[] = {
	{"GoalReachingForce", UMANS_COST_GOAL_REACHING, 0.0f, 1, {{"agentMass", 0, 1.0f, true}}, 1},
	{"PowerLaw", UMANS_COST_POWER_LAW, 5.0f, 3, {{"agentMass", 0, 1.0f, true}, {"k", 1, 1.5f, true}, {"tau0", 2, 3.0f, true}}, 3},
	{"SocialForcesAvoidance", UMANS_COST_SOCIAL_FORCES, 5.0f, 6,
	 {{"agentMass", 0, 1.0f, true}, {"dt", 1, 1.0f, true}, {"V0", 2, 2.1f, true}, {"sigma", 3, 0.3f, true},
	  {"U0", 2, 0.0f, false}, {"R", 4, 0.2f, true}}, 5},
	{"ORCA", UMANS_COST_ORCA, 5.0f, 1, {{"timeHorizon", 0, 5.0f, true}}, 1},
	{"Karamouzas", UMANS_COST_KARAMOUZAS, 5.0f, 5,
	 {{"alpha", 0, 5.0f, true}, {"beta", 1, 0.5f, true}, {"gamma", 2, 1.0f, true}, {"delta", 3, 1.0f, true},
	  {"t_max", 4, 8.0f, true}}, 5},
	{"RVO", UMANS_COST_RVO, 5.0f, 1, {{"w", 0, 1.0f, true}}, 1},
	{"Moussaid", UMANS_COST_MOUSSAID, 5.0f, 1, {{"d_max", 0, 8.0f, true}}, 1},
	{"TtcaDca", UMANS_COST_TTCA_DCA, 5.0f, 4,
	 {{"sigmaAngle_goal", 0, 2.0f, true}, {"sigmaSpeed_goal", 1, 2.0f, true}, {"sigmaTtca", 2, 1.0f, true},
	  {"sigmaDca", 3, 0.3f, true}}, 4},
	{"Paris", UMANS_COST_PARIS, 5.0f, 3, {{"w_a", 0, 0.5f, true}, {"w_b", 1, 0.0f, true}, {"t_max", 2, 8.0f, true}}, 3},
	{"VanToll", UMANS_COST_VAN_TOLL, 5.0f, 1, {{"max_distance", 0, 8.0f, true}}, 1},
};
const int kNumCosts = (int)(sizeof(kCosts) / sizeof(kCosts[0]));

const CostDef* by_kind(int kind) {
	for (const CostDef& c : kCosts)
		if (c.kind == kind) return &c;
	return nullptr;
}

} // namespace

extern "C" int umans_cost_id(const char* name) {
	if (!name) return -1;
	for (const CostDef& c : kCosts)
		if (std::strcmp(c.name, name) == 0) return c.kind;
	return -1;
}

extern "C" const char* umans_cost_name(int kind) {
	const CostDef* c = by_kind(kind);
	return c ? c->name : nullptr;
}

// Index of the parameter slot an XML attribute writes ("range" and "coeff" are not
// slots: they are CostFunction::range_ and the policy coefficient).
extern "C" int umans_cost_param_index(int kind, const char* param) {
	const CostDef* c = by_kind(kind);
	if (!c || !param) return -1;
	for (int i = 0; i < c->nparam; ++i)
		if (std::strcmp(c->params[i].name, param) == 0) return c->params[i].slot;
	return -1;
}

// Name of slot `index` (its first attribute name in read order).
extern "C" const char* umans_cost_param_name(int kind, int index) {
	const CostDef* c = by_kind(kind);
	if (!c || index < 0 || index >= c->nslots) return nullptr;
	for (int i = 0; i < c->nparam; ++i)
		if (c->params[i].slot == index) return c->params[i].name;
	return nullptr;
}

// Attribute names in the order the cost function's parseParameters reads them.
extern "C" const char* umans_cost_param_read_name(int kind, int q) {
	const CostDef* c = by_kind(kind);
	if (!c || q < 0 || q >= c->nparam) return nullptr;
	return c->params[q].name;
}

extern "C" int umans_cost_defaults(int kind, float* range, float params[UMANS_MAX_PARAMS]) {
	const CostDef* c = by_kind(kind);
	if (!c) return UMANS_EINVAL;
	if (range) *range = c->range;
	if (params) {
		for (int i = 0; i < UMANS_MAX_PARAMS; ++i) params[i] = 0.0f;
		for (int i = 0; i < c->nparam; ++i)
			if (c->params[i].has_default) params[c->params[i].slot] = c->params[i].dflt;
	}
	return UMANS_OK;
}

// Policy defaults (policy.h:63-71, 121-125).
extern "C" int umans_policy_init(umans_policy* p, int method) {
	if (!p || method < 0 || method > 2) return UMANS_EINVAL;
	std::memset(p, 0, sizeof(*p));
	p->method = method;
	p->s_type = 0;
	p->s_base = 0;
	p->s_base_dir = 1;
	p->s_radius = 0;
	p->s_angle = 180.0f;
	p->s_speed_samples = 4;
	p->s_angle_samples = 11;
	p->s_random_samples = 100;
	p->s_include_base = 0;
	p->relaxation_time = 0.0f;
	p->contact_force_scale = 5000.0f;
	p->n_cost = 0;
	p->gradient_iterations = 1;
	p->gradient_step = 1.0f;
	return UMANS_OK;
}

int umans_internal_num_cost_kinds() { return kNumCosts; }
