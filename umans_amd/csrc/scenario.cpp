// Scenario loader (host only): the reference's XML schema read into packed policies and
// agent arrays. It follows CrowdSimulator::FromConfigFile and its block loaders
// (src/Engine/core/crowdSimulator.cpp:167-647): configPath redirection, World type and
// size, mandatory delta_time, optional end_time, Policies / Agents / Obstacles blocks
// inline or in a `file=` attribute relative to the scenario's folder, policy and agent
// attributes with the same names and defaults, agent IDs assigned as
// WorldBase::AddAgent assigns them (worldBase.cpp:161-187). Attribute values are parsed
// as tinyxml2's Query*Attribute parses them (sscanf %f / %d; bools as an integer or
// "true"/"false").
#include "umans_gpu.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------ minimal XML reader
struct Node {
	std::string name;
	std::vector<std::pair<std::string, std::string>> attrs;
	std::vector<std::unique_ptr<Node>> kids;

	const char* attr(const char* n) const {
		for (const auto& a : attrs)
			if (a.first == n) return a.second.c_str();
		return nullptr;
	}
	const Node* child(const char* n = nullptr) const {
		for (const auto& k : kids)
			if (!n || k->name == n) return k.get();
		return nullptr;
	}
	// tinyxml2's NextSiblingElement(name) from the element `after`
	const Node* next_after(const Node* after, const char* n) const {
		bool seen = false;
		for (const auto& k : kids) {
			if (seen && (!n || k->name == n)) return k.get();
			if (k.get() == after) seen = true;
		}
		return nullptr;
	}
};

std::string decode_entities(const std::string& s) {
	std::string out;
	for (size_t i = 0; i < s.size(); ++i) {
		if (s[i] != '&') { out += s[i]; continue; }
		const size_t e = s.find(';', i);
		if (e == std::string::npos) { out += s[i]; continue; }
		const std::string ent = s.substr(i + 1, e - i - 1);
		if (ent == "amp") out += '&';
		else if (ent == "lt") out += '<';
		else if (ent == "gt") out += '>';
		else if (ent == "quot") out += '"';
		else if (ent == "apos") out += '\'';
		else { out += s.substr(i, e - i + 1); }
		i = e;
	}
	return out;
}

class Parser {
public:
	explicit Parser(const std::string& t) : s(t) {}
	bool parse(std::vector<std::unique_ptr<Node>>& roots, std::string& err) {
		while (true) {
			skip_misc();
			if (p >= s.size()) return true;
			if (s[p] != '<') { err = "text outside the root element"; return false; }
			std::unique_ptr<Node> n;
			if (!element(n, err)) return false;
			roots.push_back(std::move(n));
		}
	}

private:
	const std::string& s;
	size_t p = 0;

	void ws() { while (p < s.size() && isspace((unsigned char)s[p])) ++p; }
	bool starts(const char* t) const { return s.compare(p, std::strlen(t), t) == 0; }
	void skip_misc() {
		while (true) {
			ws();
			if (starts("<?")) { const size_t e = s.find("?>", p); p = e == std::string::npos ? s.size() : e + 2; }
			else if (starts("<!--")) { const size_t e = s.find("-->", p); p = e == std::string::npos ? s.size() : e + 3; }
			else if (starts("<!")) { const size_t e = s.find('>', p); p = e == std::string::npos ? s.size() : e + 1; }
			else return;
		}
	}
	std::string name_token() {
		const size_t b = p;
		while (p < s.size() && !isspace((unsigned char)s[p]) && s[p] != '>' && s[p] != '/' && s[p] != '=') ++p;
		return s.substr(b, p - b);
	}
	bool element(std::unique_ptr<Node>& out, std::string& err) {
		++p;   // '<'
		out.reset(new Node());
		out->name = name_token();
		if (out->name.empty()) { err = "empty element name"; return false; }
		while (true) {
			ws();
			if (p >= s.size()) { err = "unterminated element <" + out->name + ">"; return false; }
			if (starts("/>")) { p += 2; return true; }
			if (s[p] == '>') { ++p; break; }
			const std::string an = name_token();
			ws();
			if (p >= s.size() || s[p] != '=') { err = "attribute without value in <" + out->name + ">"; return false; }
			++p;
			ws();
			if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) { err = "unquoted attribute in <" + out->name + ">"; return false; }
			const char q = s[p++];
			const size_t e = s.find(q, p);
			if (e == std::string::npos) { err = "unterminated attribute"; return false; }
			out->attrs.emplace_back(an, decode_entities(s.substr(p, e - p)));
			p = e + 1;
		}
		// content: child elements, text (ignored), comments
		while (true) {
			const size_t lt = s.find('<', p);
			if (lt == std::string::npos) { err = "missing </" + out->name + ">"; return false; }
			p = lt;
			if (starts("<!--")) { const size_t e = s.find("-->", p); p = e == std::string::npos ? s.size() : e + 3; continue; }
			if (starts("<![CDATA[")) { const size_t e = s.find("]]>", p); p = e == std::string::npos ? s.size() : e + 3; continue; }
			if (starts("</")) {
				p += 2;
				const std::string cn = name_token();
				ws();
				if (cn != out->name || p >= s.size() || s[p] != '>') { err = "mismatched </" + cn + "> for <" + out->name + ">"; return false; }
				++p;
				return true;
			}
			if (starts("<?")) { const size_t e = s.find("?>", p); p = e == std::string::npos ? s.size() : e + 2; continue; }
			std::unique_ptr<Node> k;
			if (!element(k, err)) return false;
			out->kids.push_back(std::move(k));
		}
	}
};

struct Doc {
	std::vector<std::unique_ptr<Node>> roots;
	const Node* first(const char* n) const {
		for (const auto& r : roots)
			if (r->name == n) return r.get();
		return nullptr;
	}
};

bool load_doc(const std::string& path, Doc& d, std::string& err) {
	std::ifstream f(path, std::ios::binary);
	if (!f) { err = "cannot open " + path; return false; }
	std::stringstream ss;
	ss << f.rdbuf();
	const std::string text = ss.str();
	Parser ps(text);
	if (!ps.parse(d.roots, err)) { err = path + ": " + err; return false; }
	return true;
}

// tinyxml2 QueryFloatAttribute / QueryIntAttribute / QueryBoolAttribute
bool qfloat(const Node* n, const char* a, float& v) {
	const char* s = n->attr(a);
	float x;
	if (s && std::sscanf(s, "%f", &x) == 1) { v = x; return true; }
	return false;
}
bool qint(const Node* n, const char* a, int& v) {
	const char* s = n->attr(a);
	int x;
	if (s && std::sscanf(s, "%d", &x) == 1) { v = x; return true; }
	return false;
}
bool qbool(const Node* n, const char* a, bool& v) {
	const char* s = n->attr(a);
	if (!s) return false;
	int x;
	if (std::sscanf(s, "%d", &x) == 1) { v = x != 0; return true; }
	if (std::strcmp(s, "true") == 0) { v = true; return true; }
	if (std::strcmp(s, "false") == 0) { v = false; return true; }
	return false;
}

std::string folder_of(const std::string& path) {
	const size_t k = path.find_last_of("/\\");
	return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

} // namespace

struct umans_scenario {
	umans_scenario_info info{};
	std::vector<umans_policy> policies;
	std::vector<int> policy_xml_id;
	std::vector<int32_t> ids;
	std::vector<float> pos, goal, viewdir, radius, pref_speed, max_speed, max_accel, start_time;
	std::vector<int32_t> policy, remove_at_goal;
	std::vector<int32_t> obstacle_count;   // points per obstacle
	std::vector<float> obstacle_points;
	std::string err;
	std::vector<std::string> warnings;
};

namespace {

int policy_index(const umans_scenario& sc, int xml_id) {
	for (size_t i = 0; i < sc.policy_xml_id.size(); ++i)
		if (sc.policy_xml_id[i] == xml_id) return (int)i;
	return -1;
}

// CrowdSimulator::FromConfigFile_loadSinglePolicy (crowdSimulator.cpp:327-422)
bool load_policy(umans_scenario& sc, const Node* e) {
	int id = 0;
	qint(e, "id", id);
	const char* m = e->attr("OptimizationMethod");
	const std::string method = m ? m : "";
	int meth;
	if (method == "gradient") meth = UMANS_METHOD_GRADIENT;
	else if (method == "sampling") meth = UMANS_METHOD_SAMPLING;
	else if (method == "global") meth = UMANS_METHOD_GLOBAL;
	else { sc.err = "Policy " + std::to_string(id) + ": optimization method invalid"; return false; }
	umans_policy p;
	umans_policy_init(&p, meth);
	if (meth == UMANS_METHOD_SAMPLING) {
		qfloat(e, "SamplingAngle", p.s_angle);
		qint(e, "SpeedSamples", p.s_speed_samples);
		qint(e, "AngleSamples", p.s_angle_samples);
		qint(e, "RandomSamples", p.s_random_samples);
		bool ib = p.s_include_base != 0;
		if (qbool(e, "IncludeBaseAsSample", ib)) p.s_include_base = ib ? 1 : 0;
		const char* r;
		if ((r = e->attr("SamplingType"))) {
			if (!std::strcmp(r, "regular")) p.s_type = 0;
			else if (!std::strcmp(r, "random")) p.s_type = 1;
			else sc.warnings.push_back("Policy " + std::to_string(id) + ": sampling type invalid, using default (regular)");
		}
		if ((r = e->attr("SamplingBase"))) {
			if (!std::strcmp(r, "zero")) p.s_base = 0;
			else if (!std::strcmp(r, "current velocity")) p.s_base = 1;
			else sc.warnings.push_back("Policy " + std::to_string(id) + ": sampling base invalid, using default (zero)");
		}
		if ((r = e->attr("SamplingBaseDirection"))) {
			if (!std::strcmp(r, "unit")) p.s_base_dir = 0;
			else if (!std::strcmp(r, "current velocity")) p.s_base_dir = 1;
			else if (!std::strcmp(r, "preferred velocity")) p.s_base_dir = 2;
			else sc.warnings.push_back("Policy " + std::to_string(id) + ": sampling base direction invalid");
		}
		if ((r = e->attr("SamplingRadius"))) {
			if (!std::strcmp(r, "preferred speed")) p.s_radius = 0;
			else if (!std::strcmp(r, "maximum speed")) p.s_radius = 1;
			else if (!std::strcmp(r, "maximum acceleration")) p.s_radius = 2;
			else sc.warnings.push_back("Policy " + std::to_string(id) + ": sampling radius type invalid");
		}
	}
	qfloat(e, "RelaxationTime", p.relaxation_time);
	qfloat(e, "ContactForceScale", p.contact_force_scale);
	// the gradient method's descent extension (umans_policy::gradient_iterations); the
	// reference's loader ignores these attributes
	qint(e, "GradientIterations", p.gradient_iterations);
	qfloat(e, "GradientStepSize", p.gradient_step);
	// cost functions: the first <costfunction> child, then every following sibling
	// element (crowdSimulator.cpp:394-403 walks NextSiblingElement() without a name)
	bool started = false;
	for (const auto& k : e->kids) {
		const Node* f = k.get();
		if (!started && f->name != "costfunction") continue;
		started = true;
		const char* nm = f->attr("name");
		const int kind = umans_cost_id(nm);
		if (kind < 0) {
			// registered by the reference (costFunctionFactory.cpp:47-60) but not on this path
			static const char* kOther[] = {"PLEdestrians", "FOEAvoidance", "RandomFunction", "Generic"};
			for (const char* o : kOther)
				if (nm && !std::strcmp(nm, o)) {
					sc.err = std::string("cost function ") + nm + " is not supported by this backend";
					return false;
				}
			// CostFunctionFactory::CreateCostFunction returns null: the reference skips it
			sc.warnings.push_back(std::string("unknown cost function ") + (nm ? nm : "(none)") + ", skipped");
			continue;
		}
		if (p.n_cost >= UMANS_MAX_COSTS) { sc.err = "Policy " + std::to_string(id) + ": more than 4 cost functions"; return false; }
		const int c = p.n_cost++;
		p.cost_kind[c] = kind;
		umans_cost_defaults(kind, &p.cost_range[c], p.cost_params[c]);
		p.cost_coeff[c] = 1.0f;
		qfloat(f, "coeff", p.cost_coeff[c]);                 // Policy::AddCostFunction (policy.cpp:143-149)
		qfloat(f, "range", p.cost_range[c]);                 // CostFunction::parseParameters (costFunction.cpp:166-169)
		for (int q = 0;; ++q) {                              // each parseParameters, in its read order
			const char* pname = umans_cost_param_read_name(kind, q);
			if (!pname) break;
			qfloat(f, pname, p.cost_params[c][umans_cost_param_index(kind, pname)]);
		}
	}
	if (p.n_cost == 0) { sc.err = "Policy " + std::to_string(id) + " needs at least one cost function element"; return false; }
	if (policy_index(sc, id) >= 0) { sc.err = "Policy " + std::to_string(id) + ": ID already taken"; return false; }
	sc.policies.push_back(p);
	sc.policy_xml_id.push_back(id);
	return true;
}

struct IdAlloc {
	std::set<long long> listed;
	long long next = 0;
	// WorldBase::AddAgent (worldBase.cpp:161-187): a desired ID is refused only when an
	// agent in the list has it (GetAgent); agents scheduled for later are not in the list
	// yet, so their IDs can be given again
	long long take(long long desired, bool joins_list) {
		long long id = (desired >= 0 && !listed.count(desired)) ? desired : next;
		if (joins_list) listed.insert(id);
		if (id >= next) next = id + 1;
		return id;
	}
};

// CrowdSimulator::FromConfigFile_loadSingleAgent (crowdSimulator.cpp:424-494)
bool load_agent(umans_scenario& sc, const Node* e, IdAlloc& ids) {
	int id = -1;
	qint(e, "id", id);
	float rad = 0.24f, pref = 1.4f, vmax = 1.8f, amax = 5.0f;   // Agent::Settings (agent.h:39-57)
	bool rag = false;
	qfloat(e, "rad", rad);
	qfloat(e, "pref_speed", pref);
	qfloat(e, "max_speed", vmax);
	qfloat(e, "max_acceleration", amax);
	qbool(e, "remove_at_goal", rag);
	const Node* pe = e->child("pos");
	if (!pe) { sc.err = "Agent " + std::to_string(id) + " needs a position element"; return false; }
	float x = 0.0f, y = 0.0f;
	qfloat(pe, "x", x);
	qfloat(pe, "y", y);
	const Node* po = e->child("Policy");
	if (!po) { sc.err = "Agent " + std::to_string(id) + " needs a policy element"; return false; }
	int pid = 0;
	qint(po, "id", pid);
	const int pix = policy_index(sc, pid);
	if (pix < 0) { sc.err = "The policy with id " + std::to_string(pid) + " doesn't exist"; return false; }
	float st = 0.0f;
	qfloat(e, "start_time", st);
	float gx = x, gy = y;
	const Node* ge = e->child("goal");
	if (ge) { qfloat(ge, "x", gx); qfloat(ge, "y", gy); }
	else sc.warnings.push_back("Agent " + std::to_string(id) + " has no goal element; it will not move");
	sc.ids.push_back((int32_t)ids.take(id >= 0 ? id : -1, st <= 0.0f));   // time_ is 0 while loading
	sc.pos.push_back(x); sc.pos.push_back(y);
	sc.goal.push_back(gx); sc.goal.push_back(gy);
	// Agent::setGoal looks at the goal unless it is the position (agent.cpp:120-127)
	float vx = 0.0f, vy = 0.0f;
	if (!(gx == x && gy == y)) {
		const float dx = gx - x, dy = gy - y;
		const float m = std::sqrt(dx * dx + dy * dy);
		if (m > 0.0f) { vx = dx / m; vy = dy / m; }
		else { vx = dx; vy = dy; }          // Vector2D::normalize leaves it (vector2D.h:55-63)
	}
	sc.viewdir.push_back(vx); sc.viewdir.push_back(vy);
	sc.radius.push_back(rad);
	sc.pref_speed.push_back(pref);
	sc.max_speed.push_back(vmax);
	sc.max_accel.push_back(amax);
	sc.policy.push_back(pix);
	sc.start_time.push_back(st);
	sc.remove_at_goal.push_back(rag ? 1 : 0);
	return true;
}

// Children of a block, as the reference walks them: the first child element, then the
// following siblings with the given name.
template <typename F>
bool each_block_child(const Node* block, const char* name, F&& f) {
	const Node* e = block->child();
	while (e) {
		if (!f(e)) return false;
		e = block->next_after(e, name);
	}
	return true;
}

bool load_block(umans_scenario& sc, const Node* block, const std::string& folder, const char* root_name,
                const std::function<bool(const Node*)>& body) {
	const char* ext = block->attr("file");
	if (!ext) return body(block);
	Doc d;
	std::string err;
	if (!load_doc(folder + ext, d, err)) { sc.err = "Could not load or parse " + std::string(root_name) + " XML file: " + err; return false; }
	const Node* r = d.first(root_name);
	if (!r) { sc.err = std::string("no ") + root_name + " element in " + folder + ext; return false; }
	return body(r);
}

bool load_file(umans_scenario& sc, const std::string& path, int depth) {
	if (depth > 8) { sc.err = "configPath nesting too deep"; return false; }
	Doc d;
	std::string err;
	if (!load_doc(path, d, err)) { sc.err = "Could not load or parse XML file: " + err; return false; }
	const std::string folder = folder_of(path);
	if (const Node* cp = d.first("configPath")) {             // crowdSimulator.cpp:532-538
		const char* p = cp->attr("path");
		return load_file(sc, folder + (p ? p : ""), depth + 1);
	}
	const Node* sim = d.first("Simulation");
	if (!sim) { sc.err = "No main Simulation element in the XML file"; return false; }
	const Node* world = sim->child("World");
	if (!world) { sc.err = "No World element in the XML file"; return false; }
	// FromConfigFile_loadWorld (crowdSimulator.cpp:167-208)
	const char* type = world->attr("type");
	sc.info.world_type = 0;
	if (type && !std::strcmp(type, "Toric")) {
		float w = -1.0f, h = -1.0f;
		qfloat(world, "width", w);
		qfloat(world, "height", h);
		if (!(w > 0.0f && h > 0.0f)) { sc.err = "No valid size specified for the toric world"; return false; }
		sc.info.world_type = 1;
		sc.info.width = w;
		sc.info.height = h;
	} else if (!type || std::strcmp(type, "Infinite")) {
		sc.warnings.push_back("No valid world type specified; selecting Infinite");
	}
	float dt = -1.0f;
	qfloat(sim, "delta_time", dt);
	if (!(dt > 0.0f)) { sc.err = "No valid value for delta_time (mandatory, positive)"; return false; }
	sc.info.delta_time = dt;
	sc.info.end_time = -1.0f;
	qfloat(sim, "end_time", sc.info.end_time);
	if (const Node* pb = sim->child("Policies")) {
		if (!load_block(sc, pb, folder, "Policies", [&](const Node* b) {
			    return each_block_child(b, "Policy", [&](const Node* e) { return load_policy(sc, e); });
		    }))
			return false;
	}
	if (sc.policies.empty()) { sc.err = "A simulation needs a Policy block with at least one valid Policy element"; return false; }
	if (const Node* ab = sim->child("Agents")) {
		IdAlloc ids;
		if (!load_block(sc, ab, folder, "Agents", [&](const Node* b) {
			    return each_block_child(b, "Agent", [&](const Node* e) { return load_agent(sc, e, ids); });
		    }))
			return false;
		// Scheduled agents can share an ID (see IdAlloc). The reference then resolves every
		// neighbour with that ID to the agent listed last (GetAgent_const,
		// worldBase.cpp:51-60) and excludes them all from each other's lists by ID
		// (AgentKDTree.cpp:48-50); this backend keeps them distinct agents.
		std::map<int32_t, int> seen;
		for (int32_t v : sc.ids) ++seen[v];
		for (const auto& kv : seen)
			if (kv.second > 1)
				sc.warnings.push_back("Agent ID " + std::to_string(kv.first) + " is shared by " + std::to_string(kv.second) +
				                      " agents: the reference resolves their neighbour entries to the one listed "
				                      "last; this backend keeps them distinct");
	}
	if (const Node* ob = world->child("Obstacles")) {          // crowdSimulator.cpp:496-514
		if (!load_block(sc, ob, folder, "Obstacles", [&](const Node* b) {
			    return each_block_child(b, "Obstacle", [&](const Node* e) {
				    int cnt = 0;
				    for (const auto& k : e->kids) {
					    if (k->name != "Point") continue;
					    float x = 0.0f, y = 0.0f;
					    qfloat(k.get(), "x", x);
					    qfloat(k.get(), "y", y);
					    sc.obstacle_points.push_back(x);
					    sc.obstacle_points.push_back(y);
					    ++cnt;
				    }
				    sc.obstacle_count.push_back(cnt);
				    return true;
			    });
		    }))
			return false;
	}
	return true;
}

} // namespace

extern "C" {

int umans_scenario_load(const char* path, umans_scenario** out, char* err, int err_len) {
	if (!path || !out) return UMANS_EINVAL;
	*out = nullptr;
	umans_scenario* sc = new umans_scenario();
	const bool ok = load_file(*sc, path, 0);
	if (!ok) {
		if (err && err_len > 0) {
			std::strncpy(err, sc->err.c_str(), (size_t)err_len - 1);
			err[err_len - 1] = 0;
		}
		delete sc;
		return UMANS_EINVAL;
	}
	sc->info.n_agents = (int32_t)sc->ids.size();
	sc->info.n_policies = (int32_t)sc->policies.size();
	sc->info.n_obstacles = (int32_t)sc->obstacle_count.size();
	sc->info.n_warnings = (int32_t)sc->warnings.size();
	*out = sc;
	return UMANS_OK;
}

void umans_scenario_free(umans_scenario* sc) { delete sc; }

int umans_scenario_get_info(const umans_scenario* sc, umans_scenario_info* info) {
	if (!sc || !info) return UMANS_EINVAL;
	*info = sc->info;
	return UMANS_OK;
}

const char* umans_scenario_warning(const umans_scenario* sc, int i) {
	if (!sc || i < 0 || i >= (int)sc->warnings.size()) return nullptr;
	return sc->warnings[i].c_str();
}

int umans_scenario_get(const umans_scenario* sc, umans_policy* policies, int32_t* policy_xml_ids, umans_agents* a,
                       float* start_time, int32_t* remove_at_goal) {
	if (!sc) return UMANS_EINVAL;
	const size_t n = sc->ids.size();
	if (policies) std::memcpy(policies, sc->policies.data(), sc->policies.size() * sizeof(umans_policy));
	if (policy_xml_ids) std::memcpy(policy_xml_ids, sc->policy_xml_id.data(), sc->policy_xml_id.size() * sizeof(int32_t));
	if (a) {
		auto cp = [&](auto* dst, const auto& src) { if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0])); };
		cp(a->id, sc->ids);
		cp(a->pos, sc->pos);
		cp(a->goal, sc->goal);
		cp(a->viewdir, sc->viewdir);
		cp(a->radius, sc->radius);
		cp(a->pref_speed, sc->pref_speed);
		cp(a->max_speed, sc->max_speed);
		cp(a->max_accel, sc->max_accel);
		cp(a->policy, sc->policy);
		if (a->vel) std::memset(a->vel, 0, 2 * n * sizeof(float));          // agent.cpp:29
		if (a->world) std::memset(a->world, 0, n * sizeof(int32_t));
		if (a->rng) {
			for (size_t i = 0; i < n; ++i) {
				uint32_t r = (uint32_t)sc->ids[i] % 2147483647u;
				a->rng[i] = r ? r : 1u;
			}
		}
	}
	if (start_time) std::memcpy(start_time, sc->start_time.data(), n * sizeof(float));
	if (remove_at_goal) std::memcpy(remove_at_goal, sc->remove_at_goal.data(), n * sizeof(int32_t));
	return UMANS_OK;
}

int umans_scenario_obstacles(const umans_scenario* sc, int32_t* counts, float* points) {
	if (!sc) return UMANS_EINVAL;
	if (counts) std::memcpy(counts, sc->obstacle_count.data(), sc->obstacle_count.size() * sizeof(int32_t));
	if (points) std::memcpy(points, sc->obstacle_points.data(), sc->obstacle_points.size() * sizeof(float));
	return UMANS_OK;
}

} // extern "C"
