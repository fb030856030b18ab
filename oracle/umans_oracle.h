/* CPU oracle for the UMANS per-step agent loop — TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C++ restatement of one WorldBase::DoStep of the reference
 * (src/Engine/core/worldBase.cpp:79-128), templated on the arithmetic type:
 *   precision 0: FP32 with the reference's own precision map (the float/double
 *                mix of src/Engine/tools/vector2D.h and the cost functions);
 *   precision 1: FP64 everywhere (the "FP64 CPU oracle of the same costs").
 * Neighbour selection is identical in both modes: the reference predicate
 * (double d2 from float coordinates, strict <, r2 = double(float(r*r)),
 * src/Engine/core/AgentKDTree.cpp:35-59, nanoflann.hpp:249-253) with the
 * canonical order (d2 ascending, ties by agent ID) and the toric 5-query
 * rule (src/Engine/core/worldToric.cpp:61-79).
 *
 * Only tests/, the smoke check and bench.py's CPU-baseline leg may load it;
 * the product (umans_amd) never does.
 *
 * Parity pinning: golden vectors produced by the compiled reference
 * (oracle/_ref, tests/golden/make_golden.py) for every cost function, both
 * world types and all three optimisation methods; closed-form and invariant
 * tests in tests/test_oracle_*.py.
 */
#ifndef UMANS_ORACLE_H
#define UMANS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cost-function kinds; numbering shared by convention with include/umans_gpu.h. */
enum {
	ORC_GOAL_REACHING = 0, ORC_POWER_LAW = 1, ORC_SOCIAL_FORCES = 2, ORC_ORCA = 3,
	ORC_KARAMOUZAS = 4, ORC_RVO = 5, ORC_MOUSSAID = 6, ORC_TTCA_DCA = 7,
	ORC_PARIS = 8, ORC_VAN_TOLL = 9
};

typedef struct {
	int32_t world_type;      /* 0 infinite, 1 toric */
	float width, height;     /* toric size */
	float delta_time;
	int32_t max_neighbors;   /* cap k on the canonical list; <= 0: no cap */
	int32_t n_edges;         /* obstacle edges (Polygon2D::GetEdges order), may be 0 */
	const float* edges;      /* n_edges x (ax, ay, bx, by) */
} oracle_world;

typedef struct {
	int32_t method;          /* 0 gradient, 1 sampling, 2 global */
	int32_t s_type;          /* 0 regular, 1 random */
	int32_t s_base;          /* 0 zero, 1 current velocity */
	int32_t s_base_dir;      /* 0 unit, 1 current velocity, 2 preferred velocity */
	int32_t s_radius;        /* 0 preferred speed, 1 maximum speed, 2 maximum acceleration */
	float s_angle;
	int32_t s_speed_samples, s_angle_samples, s_random_samples, s_include_base;
	float relaxation_time, contact_force_scale;
	int32_t n_cost;
	int32_t cost_kind[4];
	float cost_coeff[4];
	float cost_range[4];
	float cost_params[4][8];
	int32_t gradient_iterations;  /* >= 2: the gradient method's descent extension (not in the
	                                 reference): v <- v - (gradient_step dt) sum coeff grad C(v),
	                                 acceleration (v_N - v) / dt; 0 or 1: policy.cpp:78-89 */
	float gradient_step;
} oracle_policy;

/* Per-agent arrays (length n; 2-vectors interleaved x,y). */
typedef struct {
	const int32_t* id;
	const float* pos;
	const float* vel;
	const float* goal;
	const float* viewdir;
	const float* radius;
	const float* pref_speed;
	const float* max_speed;
	const float* max_accel;
	const int32_t* policy;
	const uint32_t* rng;     /* minstd_rand0 state per agent (may be null if no random sampling) */
} oracle_agents_in;

typedef struct {
	float* pos;              /* new state after the step */
	float* vel;
	float* viewdir;
	float* pref_vel;         /* preferred velocity computed in the step */
	float* next_vel;         /* policy output before contact forces */
	float* contact;          /* contact force */
	uint32_t* rng;           /* advanced RNG state (may be null) */
	/* FP64 copies of the same quantities (precision 1); may be null */
	double* vel64;
	double* next_vel64;
	/* neighbour lists (may be null): counts[n]; ids/d2/seg [n * list_stride] */
	int32_t list_stride;
	int32_t* nb_count;
	int32_t* nb_id;
	double* nb_d2;
	int32_t* nb_seg;
	int32_t* nb_full;        /* uncapped count per agent (may be null) */
} oracle_agents_out;

/* One DoStep on n agents. Returns 0, or a negative value on bad input. */
int oracle_step(const oracle_world* world, const oracle_policy* policies, int n_policies,
                int n, const oracle_agents_in* in, oracle_agents_out* out,
                int precision, int nthreads);

/* Brute-force O(N^2) neighbour lists with the same predicate (for cross-checks). */
int oracle_neighbors_bruteforce(const oracle_world* world, const oracle_policy* policies,
                                int n, const oracle_agents_in* in, int list_stride,
                                int32_t* nb_count, int32_t* nb_id, double* nb_d2, int32_t* nb_seg);

/* Initial minstd_rand0 state for an agent ID, as std::default_random_engine(id). */
uint32_t oracle_rng_seed(uint32_t id);

/* Draw `count` floats U(lo,hi) from a minstd_rand0 state through
 * std::uniform_real_distribution<float>; returns the new state. */
uint32_t oracle_rng_uniform(uint32_t state, float lo, float hi, int count, float* out);

#ifdef __cplusplus
}
#endif
#endif
