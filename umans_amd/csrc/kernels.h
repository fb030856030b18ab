// Kernels of one WorldBase::DoStep (src/Engine/core/worldBase.cpp:79-128) on gfx950.
//
// Per step, on the context's stream:
//   k_fill         clear the cell counters when the key space changed (a kernel, never
//                  a memset, so the step can be captured in a graph; afterwards the
//                  scan leaves them zero)
//   k_bbox/k_grid  infinite worlds only: bounding box -> uniform cell grid
//   k_count        (class, cell) key of every agent, counted per key (atomic, no
//                  return)
//   k_scan_lookback class-major exclusive scan of the counts (one pass): each class's
//                  agents in cell order (the class kernels' work lists), and per cell
//                  over all classes by differences (the ranges the neighbour search walks)
//   k_reorder      scatter the state into cell order, classes interleaved within a
//                  cell (the rank inside a key taken by counting the key back down to
//                  zero), and fill the class lists; preferred velocity
//                  (Agent::ComputePreferredVelocity, agent.cpp:52-59)
//   k_step         per agent (one lane): k-nearest-within-radius query in canonical
//                  order, new velocity, contact forces, integration, toric wrap
// The sorted order of one step is the input order of the next, so agents that moved
// little keep their slot and the scatter stays nearly coalesced.
#pragma once
#include "costs.h"

namespace umans {

// ------------------------------------------------------------------ utilities
// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8). Each XCD is
// given a contiguous run of tiles instead, so the agents its CUs hold are neighbours in
// the sorted order and its L2 keeps the cells they gather from (the step kernels) and the
// per-cell counters and starts they update and read (the reorder).
#ifndef UMANS_XCD_REMAP
#define UMANS_XCD_REMAP 1
#endif
constexpr int kNumXcd = 8;
UD int xcd_tile(int b, int nb) {
#if UMANS_XCD_REMAP
	const int per = nb / kNumXcd, rem = nb % kNumXcd;
	const int x = b % kNumXcd, i = b / kNumXcd;
	return x * per + (x < rem ? x : rem) + i;
#else
	(void)nb;
	return b;
#endif
}
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ void k_fill(unsigned* p, int n, unsigned v) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
#endif

UD int ordered_key(float f) {
	const int i = __float_as_int(f);
	return i >= 0 ? i : (i ^ 0x7fffffff);
}
UD float from_ordered_key(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// The value of lane ^ J (every lane active): DPP for J = 1, 2 (quad permutes) and 8 (a
// rotation by 8 within rows of 16), ds_swizzle for 4 and 16 (within 32 lanes), a
// permute for 32 — the first three never leave the VALU.
template <int J>
UD unsigned xor_lane(unsigned v) {
	if constexpr (J == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
	else if constexpr (J == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);
	else if constexpr (J == 8) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true);
	else if constexpr (J == 4 || J == 16) return (unsigned)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));
	else return (unsigned)__shfl_xor((int)v, J);
}
// Minimum / maximum over the wave (every lane active), in every lane: within each row of
// 16 lanes by the exchanges above (three DPP moves and one swizzle), then the four rows'
// results read as scalars — instead of six lane permutes through the LDS crossbar.
UD int wave_min_i(int v) {
	v = min(v, (int)xor_lane<1>((unsigned)v));
	v = min(v, (int)xor_lane<2>((unsigned)v));
	v = min(v, (int)xor_lane<4>((unsigned)v));
	v = min(v, (int)xor_lane<8>((unsigned)v));
	const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
	const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
	return min(min(r0, r1), min(r2, r3));
}
// The copy of the step counters a wave updates (its wave index in the launch).
UD int stat_copy() {
	return (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kStatCopies - 1));
}
UD int wave_max_i(int v) {
	v = max(v, (int)xor_lane<1>((unsigned)v));
	v = max(v, (int)xor_lane<2>((unsigned)v));
	v = max(v, (int)xor_lane<4>((unsigned)v));
	v = max(v, (int)xor_lane<8>((unsigned)v));
	const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
	const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
	return max(max(r0, r1), max(r2, r3));
}

// Per world w, kBoxCopies copies of (ordered keys of min x, min y, max x, max y), each
// in a 128-byte line of its own, at bbox + (w * kBoxCopies + copy) * kBoxLine: atomics on
// one line are serialised, so the waves of a world spread theirs over the copies, and
// the grid set-up reduces them.
constexpr int kBoxCopies = 8;
constexpr int kBoxLine = 32;
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ void k_bbox_init(int* bbox, int nworlds) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nworlds * kBoxCopies) {
		int* b = bbox + (size_t)i * kBoxLine;
		b[0] = 0x7fffffff; b[1] = 0x7fffffff;
		b[2] = (int)0x80000000; b[3] = (int)0x80000000;
	}
}
#endif
// Bounding boxes of the infinite worlds: agent i's position added to its world's box
// (every lane of the wave takes part; lanes with i >= n add nothing). Agents of a batch
// are grouped by world, so a wave whose lanes share one world reduces first and issues
// one atomic per bound.
UD void bbox_add(const float2* pos, const int4* meta, const WorldInfo* worlds, int n, int i, int* bbox) {
	int w = -1;
	int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = (int)0x80000000, mxy = (int)0x80000000;
	if (i < n) {
		w = meta[i].y >> kPolicyBits;
		const float2 p = pos[i];
		if (worlds[w].world_type != 1 && isfinite(p.x) && isfinite(p.y)) {
			mnx = mxx = ordered_key(p.x);
			mny = mxy = ordered_key(p.y);
		} else {
			w = -1;
		}
	}
	const int w0 = __shfl(w, 0);
	const bool uniform = __all(w == w0 || w == -1);
	const int copy = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kBoxCopies - 1));
	if (uniform) {
		mnx = wave_min_i(mnx); mny = wave_min_i(mny); mxx = wave_max_i(mxx); mxy = wave_max_i(mxy);
		const int wl = wave_max_i(w);
		if ((threadIdx.x & 63) == 0 && wl >= 0) {
			int* b = bbox + ((size_t)wl * kBoxCopies + copy) * kBoxLine;
			atomicMin(&b[0], mnx); atomicMin(&b[1], mny);
			atomicMax(&b[2], mxx); atomicMax(&b[3], mxy);
		}
	} else if (w >= 0) {
		int* b = bbox + ((size_t)w * kBoxCopies + copy) * kBoxLine;
		atomicMin(&b[0], mnx); atomicMin(&b[1], mny);
		atomicMax(&b[2], mxx); atomicMax(&b[3], mxy);
	}
}
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ void k_copy_grids(const GridInfo* src, GridInfo* dst, int nworlds) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nworlds) dst[i] = src[i];
}
#endif
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ __launch_bounds__(256) void k_bbox(const float2* pos, const int4* meta, const WorldInfo* worlds,
                                             int n, int* bbox) {
	bbox_add(pos, meta, worlds, n, blockIdx.x * blockDim.x + threadIdx.x, bbox);
}
#endif
// One thread per world. Infinite worlds: a grid over the bounding box with cells of
// at least cell_size metres and at most cell_cap cells. Toric worlds: the world
// rectangle centred on the origin with cells at least cell_size wide.
// Also clears the step's counters (zero0[0..n0), zero1[0..n1)), saving their launches.
// reset_box: the boxes, once read, start again empty (the next reorder fills them).
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ void k_grid_setup(const WorldInfo* worlds, int* bbox, GridInfo* grids, int nworlds, unsigned* zero0,
                             int n0, unsigned* zero1, int n1, int reset_box) {
	const int wi = blockIdx.x * blockDim.x + threadIdx.x;
	if (wi < n0) zero0[wi] = 0u;
	if (wi < n1) zero1[wi] = 0u;
	if (wi >= nworlds) return;
	const WorldInfo W = worlds[wi];
	GridInfo r;
	r.pad = 0;
	double cs = W.cell_size > 0.0f ? (double)W.cell_size : 1.0;
	if (W.world_type == 1) {
		double nx = fmax(1.0, floor((double)W.W / cs)), ny = fmax(1.0, floor((double)W.H / cs));
		while (nx * ny > (double)W.cell_cap) {
			if (nx >= ny) nx = fmax(1.0, floor(nx / 2.0)); else ny = fmax(1.0, floor(ny / 2.0));
		}
		r.ox = -0.5f * W.W; r.oy = -0.5f * W.H;
		r.ncx = (int)nx; r.ncy = (int)ny;
		r.invx = (float)(nx / (double)W.W); r.invy = (float)(ny / (double)W.H);
	} else {
		int b[4] = {0x7fffffff, 0x7fffffff, (int)0x80000000, (int)0x80000000};
		for (int cp = 0; cp < kBoxCopies; ++cp) {
			int* bc = bbox + ((size_t)wi * kBoxCopies + cp) * kBoxLine;
			b[0] = min(b[0], bc[0]); b[1] = min(b[1], bc[1]);
			b[2] = max(b[2], bc[2]); b[3] = max(b[3], bc[3]);
			if (reset_box) {
				bc[0] = 0x7fffffff; bc[1] = 0x7fffffff;
				bc[2] = (int)0x80000000; bc[3] = (int)0x80000000;
			}
		}
		const bool empty = b[0] > b[2];
		const float x0 = empty ? 0.0f : from_ordered_key(b[0]);
		const float y0 = empty ? 0.0f : from_ordered_key(b[1]);
		const float x1 = empty ? 0.0f : from_ordered_key(b[2]);
		const float y1 = empty ? 0.0f : from_ordered_key(b[3]);
		const double w = (double)x1 - (double)x0, h = (double)y1 - (double)y0;
		double nx = floor(w / cs) + 1.0, ny = floor(h / cs) + 1.0;
		while (nx * ny > (double)W.cell_cap) {
			cs *= 1.25;
			nx = floor(w / cs) + 1.0;
			ny = floor(h / cs) + 1.0;
		}
		r.ox = x0; r.oy = y0;
		r.invx = (float)(1.0 / cs); r.invy = r.invx;
		r.ncx = (int)nx; r.ncy = (int)ny;
	}
	r.ncells = r.ncx * r.ncy;
	grids[wi] = r;
}
#endif

// Cell index along one axis: monotone non-decreasing in x, clamped into the grid
// (non-finite coordinates land in cell 0).
UD int bin_axis(float x, float o, float inv, int nc) {
	float t = floorf((x - o) * inv);
	t = fminf(fmaxf(t, 0.0f), (float)(nc - 1));
	return (int)t;
}

// Key of an agent: (class slot, world, cell). The class-major scan of these counts
// numbers each class's agents in cell order (the class kernels' work lists); the state
// itself is stored cell-major, every class of a cell together, so the neighbour search
// walks one contiguous range per row of cells whatever the classes.
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ __launch_bounds__(256) void k_count(const float2* pos, const int4* meta, int n, const WorldInfo* worlds,
                                              const GridInfo* grids, const DevPolicy* pol, const int* cls_slot,
                                              int slot_cells, unsigned* cell_count, unsigned* cell_of) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int4 m = meta[k];
	const int w = m.y >> kPolicyBits;
	const GridInfo g = grids[w];
	const float2 p = pos[k];
	const int slot = cls_slot[pol[m.y & kPolicyMask].cls];
	const int c = slot * slot_cells + worlds[w].cell_offset + bin_axis(p.y, g.oy, g.invy, g.ncy) * g.ncx +
	              bin_axis(p.x, g.ox, g.invx, g.ncx);
	cell_of[k] = (unsigned)c;
	atomicAdd(&cell_count[c], 1u);
}
#endif

// k_count's work for agent k of the next step, done by the step kernel that computed its
// new position (same key: the class slot of the launch, the world's fixed grid).
UD void count_next(const StepArgs& A, const WorldInfo& wi, int wix, int k, float2 p, int key_base) {
	const GridInfo g = A.next_grids[wix];
	const int c = key_base + wi.cell_offset + bin_axis(p.y, g.oy, g.invy, g.ncy) * g.ncx +
	              bin_axis(p.x, g.ox, g.invx, g.ncx);
	A.next_cell_of[k] = (unsigned)c;
	atomicAdd(&A.next_count[c], 1u);
}

// ---------------------------------------------------------------- scan (one pass)
constexpr int kScanThreads = 256;
constexpr int kScanItems = 64;   // 16384 counts per tile: few tiles for the look-back to walk
constexpr int kScanTile = kScanThreads * kScanItems;

UD unsigned block_exclusive_scan(unsigned v, unsigned* sh, unsigned& total) {
	// wave-level inclusive scan, then across the block's 4 waves
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	unsigned x = v;
	for (int o = 1; o < 64; o <<= 1) {
		const unsigned y = __shfl_up(x, o);
		if (lane >= o) x += y;
	}
	if (lane == 63) sh[wave] = x;
	__syncthreads();
	unsigned wofs = 0, tot = 0;
	for (int w = 0; w < kScanThreads / 64; ++w) {
		if (w < wave) wofs += sh[w];
		tot += sh[w];
	}
	__syncthreads();
	total = tot;
	return wofs + x - v;
}

// Single-pass exclusive scan (decoupled look-back, after Merrill & Garland, "Single-pass
// Parallel Prefix Scan with Decoupled Look-back", 2016). Each block takes a ticket in the
// order blocks start and scans tile `ticket`; it publishes the tile's total, then adds
// the totals of the tiles before it, walking back until one has published its inclusive
// prefix. A tile waits only on tiles with lower tickets, whose blocks have started and
// publish their totals without waiting on anyone, so the walk always ends (tile 0
// publishes its inclusive prefix at once). The wait is bounded all the same: after
// spin_limit sleeps (negative: at once) the tile stops waiting and sums the counts of
// every earlier tile itself — they are all in memory before the scan starts — so its
// prefix is exact whatever the other tiles do. status[t] = flag | value (flag kScanAgg:
// the tile's total, kScanInc: its inclusive prefix; values below 2^30); `ticket` and
// status[0..nb) must be zero at launch. Also clears the step's counters
// (zero0[0..n0), zero1[0..n1)).
constexpr unsigned kScanAgg = 1u << 30, kScanInc = 2u << 30, kScanValMask = kScanAgg - 1u;
constexpr int kScanSpinLimit = 1 << 20;   // sleeps before a tile sums its predecessors' counts itself
UD unsigned ld_status(unsigned* p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
UD void st_status(unsigned* p, unsigned v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ __launch_bounds__(kScanThreads) void k_scan_lookback(const unsigned* in, unsigned* out, unsigned* ticket,
                                                                unsigned* status, int n, int nb, unsigned* zero0,
                                                                int n0, unsigned* zero1, int n1, int spin_limit) {
	__shared__ unsigned sh[kScanThreads / 64];
	__shared__ unsigned s_tile, s_prefix;
	for (int i = blockIdx.x * kScanThreads + threadIdx.x; i < n0; i += gridDim.x * kScanThreads) zero0[i] = 0u;
	for (int i = blockIdx.x * kScanThreads + threadIdx.x; i < n1; i += gridDim.x * kScanThreads) zero1[i] = 0u;
	if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
	__syncthreads();
	const int tile = (int)s_tile;
	if (tile >= nb) return;          // one ticket per block: not reached
	const int base = tile * kScanTile + threadIdx.x * kScanItems;
	unsigned v[kScanItems];
	const bool whole = base + kScanItems <= n;
	if (whole) {
		// 64-byte aligned: four 16-byte loads
		const uint4* in4 = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
		for (int q = 0; q < kScanItems / 4; ++q) {
			const uint4 x = in4[q];
			v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
		}
	} else {
#pragma unroll
		for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : 0u;
	}
	unsigned s = 0;
#pragma unroll
	for (int i = 0; i < kScanItems; ++i) s += v[i];
	unsigned total;
	unsigned run = block_exclusive_scan(s, sh, total);
	// the first wave walks back over windows of 64 predecessors at once
	if (threadIdx.x < 64) {
		const int lane = threadIdx.x;
		unsigned prefix = 0;
		if (tile == 0) {
			if (lane == 0) st_status(&status[0], kScanInc | total);
		} else {
			if (lane == 0) st_status(&status[tile], kScanAgg | total);
			bool own_sum = spin_limit < 0;
			for (int top = tile - 1, spins = 0; !own_sum;) {
				const int j = top - lane;
				// before tile 0: an inclusive prefix of zero
				const unsigned w = j >= 0 ? ld_status(&status[j]) : kScanInc;
				const unsigned flag = w & ~kScanValMask;
				const unsigned long long inc = __ballot(flag == kScanInc);
				const unsigned long long unpublished = __ballot(flag == 0u);
				const int first_inc = inc ? __builtin_ctzll(inc) : 64;
				const unsigned long long needed = first_inc >= 63 ? ~0ull : ((2ull << first_inc) - 1ull);
				if (unpublished & needed) {
					if (++spins > spin_limit) {
						own_sum = true;
						break;
					}
					__builtin_amdgcn_s_sleep(1);
					continue;
				}
				// the window's values up to and including the nearest inclusive prefix
				unsigned val = lane <= first_inc ? (w & kScanValMask) : 0u;
				for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o);
				prefix += val;
				if (first_inc < 64) break;
				top -= 64;
			}
			if (own_sum) {
				// the counts of tiles 0 .. tile-1: tile * kScanTile of them (a multiple of 4,
				// within n), read 16 bytes at a time by the wave
				const int end = tile * kScanTile;
				unsigned val = 0;
				for (int i = 4 * lane; i < end; i += 256) {
					const uint4 x = *reinterpret_cast<const uint4*>(in + i);
					val += x.x + x.y + x.z + x.w;
				}
				for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o);
				prefix = val;
			}
			if (lane == 0) st_status(&status[tile], kScanInc | (prefix + total));
		}
		if (lane == 0) s_prefix = prefix;
	}
	__syncthreads();
	run += s_prefix;
#pragma unroll
	for (int i = 0; i < kScanItems; ++i) {
		const unsigned x = v[i];
		v[i] = run;
		run += x;
	}
	if (whole) {
		uint4* out4 = reinterpret_cast<uint4*>(out + base);
#pragma unroll
		for (int q = 0; q < kScanItems / 4; ++q) out4[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
	} else {
		for (int i = 0; i < kScanItems; ++i)
			if (base + i < n) out[base + i] = v[i];
	}
}
#endif

// --------------------------------------------------------------------- reorder
// Cell-major start of cell c from the class-major scan cs of the (slot, cell) counts:
// the agents of every slot in the cells before c, sum over q of cs[q, c] - cs[q, 0]
// (c = slot_cells gives the agent count).
UD unsigned cell_first_of(const unsigned* cs, unsigned c, int nslots, unsigned slot_cells) {
	unsigned t = 0;
	for (int q = 0; q < nslots; ++q) t += cs[q * slot_cells + c] - cs[q * slot_cells];
	return t;
}

// Thread i < n: agent i, with key (slot, cell), takes a rank r inside its key by counting
// the key down (the counters end at zero, ready for the next count; the order inside a
// key is that of the atomics, which no result depends on: neighbour lists are
// canonical). It goes to d = cell_first(cell) + (agents of the lower slots in that cell:
// differences of class_start) + r and is entry class_start[key] + r of the class lists.
// Thread i <= slot_cells (with several slots): cell_first[i] for the neighbour search.
// Threads i <= n_scan_state also clear the single-pass scan's ticket and tile statuses
// for the next step (the scan has finished: it runs before, on the same stream). With
// bbox, the agents' positions (those this step starts from) are added to the boxes of
// their infinite worlds, from which the next step's grids are set up: one step old, they
// may leave an agent outside its grid, which the clamped binning puts in an edge cell,
// so lists and results do not depend on it.
#ifndef UMANS_STEP_KERNELS_ONLY
__global__ __launch_bounds__(256) void k_reorder(State src, State dst, int n, const unsigned* cell_of,
                                                unsigned* cell_count, const unsigned* class_start, int nslots,
                                                int slot_cells, unsigned* cell_first, int* class_list,
                                                unsigned* scan_state, int n_scan_state, const WorldInfo* worlds,
                                                int* bbox) {
	const int k = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
	if (k < n_scan_state) scan_state[k] = 0u;
	if (bbox) bbox_add(src.pos, src.meta, worlds, n, k, bbox);
	const unsigned sc = (unsigned)slot_cells;
	if (nslots > 1 && k <= slot_cells) cell_first[k] = cell_first_of(class_start, (unsigned)k, nslots, sc);
	if (k >= n) return;
	const unsigned key = cell_of[k];
	const unsigned r = atomicSub(&cell_count[key], 1u) - 1u;
	const unsigned slot = key / sc, cell = key - slot * sc;
	unsigned d = r;
	if (nslots > 1) {
		d += cell_first_of(class_start, cell, nslots, sc);
		for (unsigned q = 0; q < slot; ++q) d += class_start[q * sc + cell + 1] - class_start[q * sc + cell];
	} else {
		d += class_start[key];
	}
	class_list[class_start[key] + r] = (int)d;
	const float2 p = src.pos[k];
	const float2 g = src.goal[k];
	const float4 prm = src.prm[k];
	dst.pos[d] = p;
	dst.vel[d] = src.vel[k];
	dst.goal[d] = g;
	dst.viewdir[d] = src.viewdir[k];
	dst.prm[d] = prm;
	dst.meta[d] = src.meta[k];
	// Agent::ComputePreferredVelocity + hasReachedGoal (agent.cpp:52-59, 99-102)
	const V2 toGoal = v2(g) - v2(p);
	V2 pref{0.0f, 0.0f};
	if (!(sqr_mag(toGoal) <= prm.x * prm.x)) pref = normalized(toGoal) * prm.y;
	dst.pref[d] = f2(pref);
}
#endif

// ------------------------------------------------------------ neighbour search
// All agents j with double d2 = (qx - x_j)^2 + (qy - y_j)^2 < double(float(r*r)), the
// query point q being float(position + displacement), the querying agent excluded
// (AgentKDTree.cpp:35-59, nanoflann.hpp:249-253 and :432-440). Toric worlds run one
// query per axis displacement whose condition holds, (0,0), +W, -W, +H, -H, and
// concatenate the results (worldToric.cpp:61-79). Each segment is kept in canonical
// order (d2 ascending, ties by agent ID); the list keeps the first kcap entries of the
// concatenation. Returns the list length; `total` is the uncapped length.
//
// The list holds float(d2) beside each entry (8 B per entry in LDS instead of 12, for
// occupancy). Rounding to float is monotonic, so unequal floats order the entries as
// their doubles do; equal floats are resolved by recomputing both exact doubles.

// Query displacement of segment s (0: none, 1: +W, 2: -W, 3: +H, 4: -H).
UD V2 seg_disp(const WorldInfo& wi, unsigned s) {
	return V2{s == 1 ? wi.W : (s == 2 ? -wi.W : 0.0f), s == 3 ? wi.H : (s == 4 ? -wi.H : 0.0f)};
}
// The exact double d2 of a list entry (sorted index | segment << kSegShift), as the
// search computed it.
UD double entry_d2(const State& S, V2 p, const WorldInfo& wi, unsigned u) {
	const V2 d = seg_disp(wi, u >> kSegShift);
	const V2 q{p.x + d.x, p.y + d.y};
	const float2 pj = S.pos[u & kIndexMask];
	const double ddx = (double)q.x - (double)pj.x;
	const double ddy = (double)q.y - (double)pj.y;
	return ddx * ddx + ddy * ddy;
}
// Does candidate j (d2, fd2 = float(d2)) order after entry (fe, u)?
UD bool orders_after(const State& S, V2 p, const WorldInfo& wi, double d2, float fd2, int j, float fe, unsigned u) {
	if (fd2 != fe) return fd2 > fe;
	const double de = entry_d2(S, p, wi, u);
	if (d2 != de) return d2 > de;
	return S.meta[j].x > S.meta[u & kIndexMask].x;
}

// entry (index | segment) u before entry w in the canonical order: segment, exact d2, ID
UD bool entry_before(const State& S, V2 p, const WorldInfo& wi, unsigned u, unsigned w) {
	if ((u >> kSegShift) != (w >> kSegShift)) return (u >> kSegShift) < (w >> kSegShift);
	const double du = entry_d2(S, p, wi, u), dw = entry_d2(S, p, wi, w);
	if (du != dw) return du < dw;
	return S.meta[u & kIndexMask].x < S.meta[w & kIndexMask].x;
}

template <int KMAX, int BS>
UD int find_neighbors(const StepArgs& A, const WorldInfo& wi, const GridInfo& g, int self, V2 p, float range,
                      float* d2s, unsigned* nbs, int& total) {
	const State& S = A.in;
	const float r2f = range * range;
	const double r2 = (double)r2f;
	const float rs = range * 1.0001f + 1e-3f;
	const int kcap = A.kcap < KMAX ? A.kcap : KMAX;   // a tier's list holds at most KMAX
	int filled = 0;
	total = 0;
	for (int s = 0; s < 5; ++s) {
		float dx = 0.0f, dy = 0.0f;
		if (s > 0) {
			if (wi.world_type != 1) break;
			if (s == 1) { if (!((double)(p.x - range) < -0.5 * (double)wi.W)) continue; dx = wi.W; }
			else if (s == 2) { if (!((double)(p.x + range) > 0.5 * (double)wi.W)) continue; dx = -wi.W; }
			else if (s == 3) { if (!((double)(p.y - range) < -0.5 * (double)wi.H)) continue; dy = wi.H; }
			else { if (!((double)(p.y + range) > 0.5 * (double)wi.H)) continue; dy = -wi.H; }
		}
		const V2 q{p.x + dx, p.y + dy};
		// cell range: a margin covers the rounding of r2 and of the float bounds
		const float m = rs + (fabsf(q.x) + fabsf(q.y)) * 1e-6f;
		const int cx0 = bin_axis(q.x - m, g.ox, g.invx, g.ncx), cx1 = bin_axis(q.x + m, g.ox, g.invx, g.ncx);
		const int cy0 = bin_axis(q.y - m, g.oy, g.invy, g.ncy), cy1 = bin_axis(q.y + m, g.oy, g.invy, g.ncy);
		const double qx = (double)q.x, qy = (double)q.y;
		const int seg_base = filled;
		const int cap = kcap - filled;
		int seg_cnt = 0;
		for (int cy = cy0; cy <= cy1; ++cy) {
			const int row = wi.cell_offset + cy * g.ncx;
			const unsigned b = A.cell_start[row + cx0], e = A.cell_start[row + cx1 + 1];
			for (unsigned j = b; j < e; ++j) {
				// one divergent branch per candidate: the exact test costs a few FP64 ops
				// (full rate) where nested early-outs cost scalar exec-mask bookkeeping
				const float2 pj = S.pos[j];
				const double ddx = qx - (double)pj.x;
				const double ddy = qy - (double)pj.y;
				const double d2 = ddx * ddx + ddy * ddy;
				const bool hit = d2 < r2 && (int)j != self;
				total += hit ? 1 : 0;
				if (!hit || cap <= 0) continue;
				const float fd2 = (float)d2;
				int at;
				if (seg_cnt < cap) {
					at = seg_base + seg_cnt;
					++seg_cnt;
				} else {
					const int last = seg_base + cap - 1;
					if (orders_after(S, p, wi, d2, fd2, (int)j, d2s[last * BS], nbs[last * BS])) continue;
					at = last;
				}
				while (at > seg_base) {
					const float fp = d2s[(at - 1) * BS];
					const unsigned np = nbs[(at - 1) * BS];
					if (orders_after(S, p, wi, d2, fd2, (int)j, fp, np)) break;
					d2s[at * BS] = fp;
					nbs[at * BS] = np;
					--at;
				}
				d2s[at * BS] = fd2;
				nbs[at * BS] = j | ((unsigned)s << kSegShift);
			}
		}
		filled += seg_cnt;
	}
	return filled;
}

// A long-list entry packed into one 64-bit key: segment (bits 61-59), the bits of
// float(d2) (58-28; d2 >= 0, so the sign bit is clear) and the sorted index (27-0).
// Rounding to float is monotonic, so the keys order entries as (segment, double d2)
// does wherever their floats differ; two entries of one row with equal segment and
// float are resolved by their exact doubles and agent IDs.
constexpr unsigned long long kKeyPad = ~0ull;   // padding: orders after every entry
UD unsigned long long long_key(int s, double d2, unsigned j) {
	return ((unsigned long long)s << 59) | ((unsigned long long)__float_as_uint((float)d2) << 28) | j;
}
UD unsigned key_entry(unsigned long long k) {   // sorted index | segment << kSegShift
	return (unsigned)(k & kIndexMask) | ((unsigned)(k >> 59) << kSegShift);
}

// Long lists, step 1 (one lane per agent): the same queries and exact test as
// find_neighbors, but every candidate in range is appended, unsorted, to the lane's row of
// the collection buffer (packed keys, below; a row is kLongCap contiguous keys, so a lane
// fills whole lines of its own and a row is read back in one coalesced access), at most
// kLongCap of them.
// Returns the number in range (all of them, as find_neighbors' total).
UD int collect_neighbors(const StepArgs& A, const WorldInfo& wi, const GridInfo& g, int self, V2 p, float range,
                         int row) {
	const State& S = A.in;
	const double r2 = (double)(range * range);
	const float rs = range * 1.0001f + 1e-3f;
	int n = 0;
	for (int s = 0; s < 5; ++s) {
		float dx = 0.0f, dy = 0.0f;
		if (s > 0) {
			if (wi.world_type != 1) break;
			if (s == 1) { if (!((double)(p.x - range) < -0.5 * (double)wi.W)) continue; dx = wi.W; }
			else if (s == 2) { if (!((double)(p.x + range) > 0.5 * (double)wi.W)) continue; dx = -wi.W; }
			else if (s == 3) { if (!((double)(p.y - range) < -0.5 * (double)wi.H)) continue; dy = wi.H; }
			else { if (!((double)(p.y + range) > 0.5 * (double)wi.H)) continue; dy = -wi.H; }
		}
		const V2 q{p.x + dx, p.y + dy};
		const float m = rs + (fabsf(q.x) + fabsf(q.y)) * 1e-6f;
		const int cx0 = bin_axis(q.x - m, g.ox, g.invx, g.ncx), cx1 = bin_axis(q.x + m, g.ox, g.invx, g.ncx);
		const int cy0 = bin_axis(q.y - m, g.oy, g.invy, g.ncy), cy1 = bin_axis(q.y + m, g.oy, g.invy, g.ncy);
		const double qx = (double)q.x, qy = (double)q.y;
		for (int cy = cy0; cy <= cy1; ++cy) {
			const int crow = wi.cell_offset + cy * g.ncx;
			const unsigned b = A.cell_start[crow + cx0], e = A.cell_start[crow + cx1 + 1];
			// two candidates a pass, both positions loaded first: one wait covers both
			auto visit = [&](unsigned j, float2 pj) {
				const double ddx = qx - (double)pj.x;
				const double ddy = qy - (double)pj.y;
				const double d2 = ddx * ddx + ddy * ddy;
				if (!(d2 < r2) || (int)j == self) return;
				if (n < kLongCap) A.long_key[(size_t)row * kLongCap + n] = long_key(s, d2, j);
				++n;
			};
			unsigned j = b;
			for (; j + 1 < e; j += 2) {
				const float2 p0 = S.pos[j], p1 = S.pos[j + 1];
				visit(j, p0);
				visit(j + 1, p1);
			}
			if (j < e) visit(j, S.pos[j]);
		}
	}
	return n;
}

// Long lists, step 2 (the whole wave, in uniform control flow): the wave sorts its rows
// with bitonic networks across lanes, in the canonical order (segment, d2, agent ID),
// and writes each row's first kcap entries into long_u in that order. Rows of up to 16,
// 32, 64 and 128 entries are sorted eight, four, two and one at a time, two entries per
// lane (rank_rows). The networks sort the packed keys;
// a pass in which two neighbouring entries of a row share segment and float(d2) sorts
// its rows again by (segment, exact d2, agent ID), the doubles recomputed from the
// positions as the search computed them. n_row: the row's entry count (0 for rows the
// wave does not sort); p, wi: the lane's own agent position and world.
UD unsigned long long shfl_xor_u64(unsigned long long v, int m) {
	const int lo = __shfl_xor((int)v, m), hi = __shfl_xor((int)(v >> 32), m);
	return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
UD unsigned long long shfl_u64(unsigned long long v, int src) {
	const int lo = __shfl((int)v, src), hi = __shfl((int)(v >> 32), src);
	return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
UD double shfl_xor_d(double v, int m) {
	return __longlong_as_double((long long)shfl_xor_u64((unsigned long long)__double_as_longlong(v), m));
}
UD bool key_less(double da, unsigned ua, int ia, double db, unsigned ub, int ib) {
	const unsigned sa = ua >> kSegShift, sb = ub >> kSegShift;
	if (sa != sb) return sa < sb;
	if (da != db) return da < db;
	return ia < ib;
}
UD bool same_float_key(unsigned long long a, unsigned long long b) {
	return a != kKeyPad && ((a ^ b) >> 28) == 0;
}

// ------------------------------------------------------------------ short lists
// Lists of up to kShortList entries (caps of 12 and 16): every candidate in range is
// appended, unsorted, to the lane's column of a small LDS buffer (one store), and the
// buffer is sorted once at the end by a fixed compare-exchange network in registers;
// a buffer that fills up drops its canonical last entry. Sorting the packed
// keys (segment, float(d2), index) orders the entries canonically wherever their floats
// differ; runs of equal segment and float are put in exact-d2 and ID order before
// anything is cut. This replaces an insertion per candidate, whose shift loop ran as
// long as the wave's longest shift and cost more instructions than the rest of the step.
// The buffer holds KMAX + 4 keys, at least MIN: a lane whose buffer fills drops its
// canonical last entry, a scan of the whole buffer, so a larger buffer drops less often at
// the price of a longer sort network and more LDS. The step kernel (one agent per lane)
// measured fastest with 20 keys for caps of both 12 and 16 (a crowd with ~12.6 agents in
// range: 24 keys at cap 16 was slower than 20, 16 at cap 12 too); the cooperative kernel
// keeps KMAX + 4.
constexpr int kShortList = 16;
constexpr int kStepMinKeys = 20, kCoopMinKeys = 0;
template <int KMAX, int MIN> constexpr int short_buffer() { return KMAX + 4 > MIN ? KMAX + 4 : MIN; }

// Batcher's odd-even merge sort for n keys: the network for the next power of two with
// the comparators that touch a slot >= n left out (those slots would hold +inf).
template <int N>
struct SortNet {
	int a[160], b[160], n;
	constexpr SortNet() : a(), b(), n(0) {
		int P = 1;
		while (P < N) P <<= 1;
		for (int q = 1; q < P; q <<= 1)
			for (int k = q; k >= 1; k >>= 1)
				for (int j = k % q; j + k < P; j += 2 * k)
					for (int i = 0; i < k; ++i)
						if (i + j + k < N && (i + j) / (2 * q) == (i + j + k) / (2 * q)) {
							a[n] = i + j;
							b[n] = i + j + k;
							++n;
						}
	}
};
template <int N> inline constexpr SortNet<N> kSortNet{};

template <int N, int C = 0>
UD void sort_net(unsigned long long* k) {
	if constexpr (C < kSortNet<N>.n) {
		constexpr int x = kSortNet<N>.a[C], y = kSortNet<N>.b[C];
		const unsigned long long lo = k[x] < k[y] ? k[x] : k[y];
		const unsigned long long hi = k[x] < k[y] ? k[y] : k[x];
		k[x] = lo;
		k[y] = hi;
		sort_net<N, C + 1>(k);
	}
}

// A full buffer drops its canonical last entry: KC - 1 >= kcap entries order before it,
// so it is in no list, nor is any later candidate whose (segment, float) is above its
// own. Returns the dropped key; the caller's count goes down by one.
template <int KC, int CS = 64>
UD unsigned long long drop_last(const State& S, V2 p, const WorldInfo& wi, unsigned long long* buf) {
	int mi = 0;
	unsigned long long mk = buf[0];
#pragma unroll
	for (int i = 1; i < KC; ++i) {
		const unsigned long long v = buf[i * CS];
		// (segment, float) decides; equal ones by exact d2 and ID (rare)
		bool later = (v >> 28) > (mk >> 28);
		if (same_float_key(v, mk)) later = entry_before(S, p, wi, key_entry(mk), key_entry(v));
		mi = later ? i : mi;
		mk = later ? v : mk;
	}
	buf[mi * CS] = buf[(KC - 1) * CS];
	return mk;
}

// Sorts the lane's cnt (<= KC) buffered keys (buf[i * CS]) into k[0..KC), pads after
// them, puts runs of equal segment and float that reach into the kept entries in exact
// d2 and ID order, and returns min(cnt, kcap).
template <int KC, int CS = 64>
UD int sort_short(const State& S, V2 p, const WorldInfo& wi, unsigned long long* buf, int cnt, int kcap,
                  unsigned long long* k) {
#pragma unroll
	for (int i = 0; i < KC; ++i) {
		const unsigned long long v = buf[i * CS];
		k[i] = i < cnt ? v : kKeyPad;
	}
	sort_net<KC>(k);
	bool tie = false;
#pragma unroll
	for (int i = 0; i + 1 < KC; ++i) tie = tie | ((i < kcap) & same_float_key(k[i], k[i + 1]));
	if (__ballot(tie) != 0ull) {
		if (tie) {
#pragma unroll
			for (int i = 0; i < KC; ++i)
				if (i < cnt) buf[i * CS] = k[i];
			for (int i = 0; i + 1 < cnt;) {
				int r = i + 1;
				while (r < cnt && same_float_key(buf[i * CS], buf[r * CS])) ++r;
				for (int x = i + 1; x < r; ++x) {
					const unsigned long long v = buf[x * CS];
					int y = x;
					while (y > i && entry_before(S, p, wi, key_entry(v), key_entry(buf[(y - 1) * CS]))) {
						buf[y * CS] = buf[(y - 1) * CS];
						--y;
					}
					buf[y * CS] = v;
				}
				i = r;
			}
#pragma unroll
			for (int i = 0; i < KC; ++i) {
				const unsigned long long v = buf[i * CS];
				k[i] = i < cnt ? v : kKeyPad;
			}
		}
	}
	return cnt < kcap ? cnt : kcap;
}

// find_neighbors for KMAX <= kShortList: the same queries, exact test and canonical list.
// buf: the lane's column of short_buffer<KMAX, MIN>() keys (stride CS, the columns of
// the wave's lanes side by side); the finished list (index | segment, stride CS) is
// written over the start of the wave's region, lane by lane as the keys were.
template <int KMAX, int MIN, int CS = 64>
UD int find_short(const StepArgs& A, const WorldInfo& wi, const GridInfo& g, int self, V2 p, float range,
                  unsigned long long* buf, unsigned* list, int& total) {
	constexpr int KC = short_buffer<KMAX, MIN>();
	const State& S = A.in;
	const double r2 = (double)(range * range);
	const float rs = range * 1.0001f + 1e-3f;
	const int kcap = A.kcap < KMAX ? A.kcap : KMAX;
	unsigned long long k[KC];
	// (segment, float) of the last dropped key (key >> 28): a candidate above it is in no list
	unsigned long long limhi = kKeyPad >> 28;
	int cnt = 0;
	total = 0;
	for (int s = 0; s < 5; ++s) {
		float dx = 0.0f, dy = 0.0f;
		if (s > 0) {
			if (wi.world_type != 1) break;
			if (s == 1) { if (!((double)(p.x - range) < -0.5 * (double)wi.W)) continue; dx = wi.W; }
			else if (s == 2) { if (!((double)(p.x + range) > 0.5 * (double)wi.W)) continue; dx = -wi.W; }
			else if (s == 3) { if (!((double)(p.y - range) < -0.5 * (double)wi.H)) continue; dy = wi.H; }
			else { if (!((double)(p.y + range) > 0.5 * (double)wi.H)) continue; dy = -wi.H; }
		}
		const V2 q{p.x + dx, p.y + dy};
		const float m = rs + (fabsf(q.x) + fabsf(q.y)) * 1e-6f;
		const int cx0 = bin_axis(q.x - m, g.ox, g.invx, g.ncx), cx1 = bin_axis(q.x + m, g.ox, g.invx, g.ncx);
		const int cy0 = bin_axis(q.y - m, g.oy, g.invy, g.ncy), cy1 = bin_axis(q.y + m, g.oy, g.invy, g.ncy);
		const double qx = (double)q.x, qy = (double)q.y;
		for (int cy = cy0; cy <= cy1; ++cy) {
			const int row = wi.cell_offset + cy * g.ncx;
			const unsigned b = A.cell_start[row + cx0], e = A.cell_start[row + cx1 + 1];
			// two candidates a pass, both positions loaded first: one wait covers both
			auto visit = [&](unsigned j, float2 pj) {
				const double ddx = qx - (double)pj.x;
				const double ddy = qy - (double)pj.y;
				const double d2 = ddx * ddx + ddy * ddy;
				const bool hit = d2 < r2 && (int)j != self;
				total += hit ? 1 : 0;
				const unsigned long long key = long_key(s, d2, j);
				if (hit && (key >> 28) <= limhi) {
					buf[cnt * CS] = key;
					if (++cnt == KC) {
						limhi = drop_last<KC, CS>(S, p, wi, buf) >> 28;
						--cnt;
					}
				}
			};
			unsigned j = b;
			for (; j + 1 < e; j += 2) {
				const float2 p0 = S.pos[j], p1 = S.pos[j + 1];
				visit(j, p0);
				visit(j + 1, p1);
			}
			if (j < e) visit(j, S.pos[j]);
		}
	}
	cnt = sort_short<KC, CS>(S, p, wi, buf, cnt, kcap, k);
	// the list overlays the keys of the wave's lanes: every lane has read its keys
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
	for (int i = 0; i < KMAX; ++i) list[i * CS] = key_entry(k[i]);
	return cnt;
}
template <int J>
UD unsigned long long xor_lane_u64(unsigned long long v) {
	return ((unsigned long long)xor_lane<J>((unsigned)(v >> 32)) << 32) | xor_lane<J>((unsigned)v);
}
// Bitonic sort, ascending, of rows of SZ entries held E per lane: entry x of a row is
// element q of lane l of the row's group of L = SZ / E lanes, x = q * L + l. Stage
// (W, J) of the merge of blocks of W compares x with x ^ J: inside the lane (element
// q ^ J / L) for J >= L, with lane l ^ J otherwise. Keys: the packed 64-bit keys.
template <int SZ, int E, int W, int J>
UD void rank_stage(unsigned long long (&k)[E], int l) {
	constexpr int L = SZ / E;
	if constexpr (J >= L) {
		constexpr int D = J / L;
#pragma unroll
		for (int q = 0; q < E; ++q) {
			if (q & D) continue;
			const bool asc = ((q * L + l) & W) == 0;
			const unsigned long long lo = k[q] < k[q + D] ? k[q] : k[q + D];
			const unsigned long long hi = k[q] < k[q + D] ? k[q + D] : k[q];
			k[q] = asc ? lo : hi;
			k[q + D] = asc ? hi : lo;
		}
	} else {
#pragma unroll
		for (int q = 0; q < E; ++q) {
			const unsigned long long kp = xor_lane_u64<J>(k[q]);
			const int x = q * L + l;
			const bool take_min = ((x & W) == 0) == ((x & J) == 0);
			k[q] = (take_min == (k[q] < kp)) ? k[q] : kp;
		}
	}
	if constexpr (J > 1) rank_stage<SZ, E, W, J / 2>(k, l);
	else if constexpr (W < SZ) rank_stage<SZ, E, 2 * W, W>(k, l);
}
// The same network on full entries (exact d2, index | segment, ID): the canonical order.
template <int SZ, int E, int W, int J>
UD void rank_exact_stage(double (&d)[E], unsigned (&u)[E], int (&id)[E], int l) {
	constexpr int L = SZ / E;
	if constexpr (J >= L) {
		constexpr int D = J / L;
#pragma unroll
		for (int q = 0; q < E; ++q) {
			if (q & D) continue;
			const bool asc = ((q * L + l) & W) == 0;
			const bool sw = asc ? key_less(d[q + D], u[q + D], id[q + D], d[q], u[q], id[q])
			                    : key_less(d[q], u[q], id[q], d[q + D], u[q + D], id[q + D]);
			if (sw) {
				const double td = d[q]; d[q] = d[q + D]; d[q + D] = td;
				const unsigned tu = u[q]; u[q] = u[q + D]; u[q + D] = tu;
				const int ti = id[q]; id[q] = id[q + D]; id[q + D] = ti;
			}
		}
	} else {
#pragma unroll
		for (int q = 0; q < E; ++q) {
			const double dp = __longlong_as_double((long long)xor_lane_u64<J>((unsigned long long)__double_as_longlong(d[q])));
			const unsigned up = xor_lane<J>(u[q]);
			const int ip = (int)xor_lane<J>((unsigned)id[q]);
			const int x = q * L + l;
			const bool take_min = ((x & W) == 0) == ((x & J) == 0);
			const bool take = take_min ? key_less(dp, up, ip, d[q], u[q], id[q]) : key_less(d[q], u[q], id[q], dp, up, ip);
			d[q] = take ? dp : d[q];
			u[q] = take ? up : u[q];
			id[q] = take ? ip : id[q];
		}
	}
	if constexpr (J > 1) rank_exact_stage<SZ, E, W, J / 2>(d, u, id, l);
	else if constexpr (W < SZ) rank_exact_stage<SZ, E, 2 * W, W>(d, u, id, l);
}

// Full entry of key k of the row whose agent stands at (px, py) in a world of width W
// and height H: exact d2 (as entry_d2 computes it), index | segment, agent ID.
UD void exact_entry(const State& S, float px, float py, float W, float H, unsigned long long k,
                    double& d, unsigned& u, int& id) {
	d = __longlong_as_double(0x7ff0000000000000ll);
	u = ~0u;
	id = 0;
	if (k == kKeyPad) return;
	u = key_entry(k);
	const unsigned sg = u >> kSegShift;
	const float qx = px + (sg == 1 ? W : (sg == 2 ? -W : 0.0f));
	const float qy = py + (sg == 3 ? H : (sg == 4 ? -H : 0.0f));
	const float2 pj = S.pos[u & kIndexMask];
	const double ddx = (double)qx - (double)pj.x;
	const double ddy = (double)qy - (double)pj.y;
	d = ddx * ddx + ddy * ddy;
	id = S.meta[u & kIndexMask].x;
}

// The rows of more than 2^(LG-1) (LG = 4: any) and at most 2^LG entries, E entries per
// lane, so 64 E >> LG rows per pass: a wave ranks its 64 rows in few passes, and a pass's
// dependent chain of cross-lane stages serves several rows. The keys of the next
// kRankAhead passes are loaded before a pass sorts, so their latency is spent sorting.
constexpr int kRankAhead = 2;
template <int LG, int E>
UD void rank_rows(const StepArgs& A, int row, int n_row, int kcap, V2 p, const WorldInfo& wi, int lane) {
	constexpr int SZ = 1 << LG, L = SZ / E, G = 64 / L;
	static_assert(L >= 1 && L <= 64 && G * L == 64, "rows of SZ entries, E per lane");
	const size_t st = (size_t)A.long_stride;
	const State& S = A.in;
	unsigned long long m = __ballot(n_row > (LG == 4 ? 0 : SZ / 2) && n_row <= SZ);
	if (!m) return;
	const int l = lane & (L - 1);
	// a pass's rows: the next G set bits of m (uniform), lane group g = lane / L; its
	// keys: entries q L + l of the lane group's row (kKeyPad where there is none)
	auto select = [&](int& src, int& n, int& rw, unsigned long long (&key)[E]) {
		int rsel = -1;
		for (int g = 0; g < G; ++g) {
			const int r = m ? __builtin_ctzll(m) : -1;
			if (m) m &= m - 1;
			if (lane / L == g) rsel = r;
		}
		src = rsel < 0 ? 0 : rsel;          // every lane shuffles
		const int n_src = __shfl(n_row, src);
		rw = __shfl(row, src);
		n = rsel >= 0 ? n_src : 0;
#pragma unroll
		for (int q = 0; q < E; ++q) {
			const int x = q * L + l;
			key[q] = x < n ? A.long_key[(size_t)rw * kLongCap + x] : kKeyPad;
		}
	};
	// the passes in flight: slot 0 is sorted next (uniform flags)
	int qsrc[kRankAhead], qn[kRankAhead], qrw[kRankAhead];
	unsigned long long qk[kRankAhead][E];
	bool qhave[kRankAhead];
#pragma unroll
	for (int d = 0; d < kRankAhead; ++d) {
		qhave[d] = m != 0ull;
		qsrc[d] = 0; qn[d] = 0; qrw[d] = 0;
#pragma unroll
		for (int q = 0; q < E; ++q) qk[d][q] = kKeyPad;
		if (qhave[d]) select(qsrc[d], qn[d], qrw[d], qk[d]);
	}
	while (qhave[0]) {
		const int csrc = qsrc[0], cn = qn[0], crw = qrw[0];
		unsigned long long k[E];
#pragma unroll
		for (int q = 0; q < E; ++q) k[q] = qk[0][q];
#pragma unroll
		for (int d = 0; d + 1 < kRankAhead; ++d) {
			qsrc[d] = qsrc[d + 1]; qn[d] = qn[d + 1]; qrw[d] = qrw[d + 1]; qhave[d] = qhave[d + 1];
#pragma unroll
			for (int q = 0; q < E; ++q) qk[d][q] = qk[d + 1][q];
		}
		qhave[kRankAhead - 1] = m != 0ull;
		if (qhave[kRankAhead - 1])
			select(qsrc[kRankAhead - 1], qn[kRankAhead - 1], qrw[kRankAhead - 1], qk[kRankAhead - 1]);
		rank_stage<SZ, E, 2, 1>(k, l);
		// ties sit next to each other once sorted: the entry after x = q L + l is element q
		// of lane l + 1, or for l = L - 1 element q + 1 of the group's first lane
		const int keep = cn < kcap ? cn : kcap;
		bool tie = false;
#pragma unroll
		for (int q = 0; q < E; ++q) {
			const unsigned long long a = shfl_u64(k[q], (lane + 1) & 63);
			const unsigned long long b = q + 1 < E ? shfl_u64(k[q + 1 < E ? q + 1 : q], lane - l) : kKeyPad;
			const unsigned long long nx = l < L - 1 ? a : b;
			tie = tie || (q * L + l + 1 < SZ && same_float_key(k[q], nx));
		}
		if (__ballot(tie)) {
			const float px = __shfl(p.x, csrc), py = __shfl(p.y, csrc);
			const float W = __shfl(wi.W, csrc), H = __shfl(wi.H, csrc);
			double d[E];
			unsigned u[E];
			int id[E];
#pragma unroll
			for (int q = 0; q < E; ++q) exact_entry(S, px, py, W, H, k[q], d[q], u[q], id[q]);
			rank_exact_stage<SZ, E, 2, 1>(d, u, id, l);
#pragma unroll
			for (int q = 0; q < E; ++q)
				if (q * L + l < keep) A.long_u[(size_t)(q * L + l) * st + crw] = u[q];
		} else {
#pragma unroll
			for (int q = 0; q < E; ++q)
				if (q * L + l < keep) A.long_u[(size_t)(q * L + l) * st + crw] = key_entry(k[q]);
		}
	}
}

UD void rank_long_lists(const StepArgs& A, int row, int n_row, int kcap, V2 p, const WorldInfo& wi) {
	const int lane = threadIdx.x & 63;
	// the rows were written by their own lanes: make those stores visible to the wave
	// (one wave per workgroup, so workgroup scope is the wave)
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
	// rows of up to 16, 32, 64 and 128 entries, two entries per lane: 8, 4, 2 and 1 per pass
	rank_rows<4, 2>(A, row, n_row, kcap, p, wi, lane);
	rank_rows<5, 2>(A, row, n_row, kcap, p, wi, lane);
	rank_rows<6, 2>(A, row, n_row, kcap, p, wi, lane);
	rank_rows<7, 2>(A, row, n_row, kcap, p, wi, lane);
	// the sorted rows are read back by their own lanes
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

// WorldBase::computeNeighboringObstacles_Flat (worldBase.cpp:62-77) for each query of
// WorldInfinite / WorldToric::ComputeNeighbors (worldToric.cpp:32-79): the edges whose
// distance to float(position + displacement) is within the range, in edge order, moved
// back by the displacement. Returns the number kept (at most cap); `total` counts them all.
UD int find_obstacles(const StepArgs& A, const WorldInfo& wi, V2 p, float range, float4* out, int& total,
                      int cap = kMaxObstacleEdges) {
	const float r2 = range * range;
	int n = 0;
	total = 0;
	for (int s = 0; s < 5; ++s) {
		float dx = 0.0f, dy = 0.0f;
		if (s > 0) {
			if (wi.world_type != 1) break;
			if (s == 1) { if (!((double)(p.x - range) < -0.5 * (double)wi.W)) continue; dx = wi.W; }
			else if (s == 2) { if (!((double)(p.x + range) > 0.5 * (double)wi.W)) continue; dx = -wi.W; }
			else if (s == 3) { if (!((double)(p.y - range) < -0.5 * (double)wi.H)) continue; dy = wi.H; }
			else { if (!((double)(p.y + range) > 0.5 * (double)wi.H)) continue; dy = -wi.H; }
		}
		const V2 q{p.x + dx, p.y + dy};
		for (int e = 0; e < A.n_edges; ++e) {
			const float4 E = A.edges[e];
			if (!(dist_line_sq(q, seg_a(E), seg_b(E), true) <= r2)) continue;
			++total;
			if (n < cap) out[n++] = make_float4(E.x - dx, E.y - dy, E.z - dx, E.w - dy);
		}
	}
	return n;
}

// --------------------------------------------------------------- per-agent state
// Policy classes: agents are bucketed by class each step and every class runs its own
// instance of the step kernel with only its code path compiled in.
// CLS_COOP: regular sampling over costs that read the neighbour list per candidate
// (Karamouzas, RVO, Moussaid, TtcaDca, Paris, VanToll): the candidates of a wave's agents
// are spread over its lanes (k_step_coop).
enum { CLS_FORCE_GRAD = 0, CLS_FORCE_SAMPLE = 1, CLS_ORCA = 2, CLS_GENERIC = 3, CLS_COOP = 4, NUM_CLS = 5 };

// Loop over a policy's cost functions: fully unrolled over the 4 slots for the force
// classes (so per-cost values stay in registers), a plain loop otherwise.
template <int CLS, typename F>
UD void for_costs(const DevPolicy& P, F&& f) {
	if constexpr (CLS == CLS_GENERIC) {
		for (int c = 0; c < P.n_cost; ++c) f(c);
	} else {
#pragma unroll
		for (int c = 0; c < 4; ++c)
			if (c < P.n_cost) f(c);
	}
}

template <int KMAX, int CLS>
struct Prep {
	V2 target[4];            // force-based costs: target velocity
	float ttcp[4];           // Karamouzas: TTC at the preferred velocity
	Lines lines;             // ORCA half-planes
	int n_lines;
};

// CostFunction::GetCost of cost c at velocity v
template <int KMAX, int CLS>
UD float cost_of(const DevPolicy& P, int c, const Self& a, const NbList& L, const Prep<KMAX, CLS>& pr, V2 v) {
	const int kind = P.kind[c];
	if constexpr (CLS == CLS_FORCE_GRAD || CLS == CLS_FORCE_SAMPLE) {
		return 0.5f * sqr_mag(v - pr.target[c]) / a.dt;       // ForceBasedFunction.cpp:26-30
	} else {
		const float* prm = P.params[c];
		const float range = P.range[c];
		switch (kind) {
		case KIND_GOAL_REACHING: case KIND_POWER_LAW: case KIND_SOCIAL_FORCES:
			return 0.5f * sqr_mag(v - pr.target[c]) / a.dt;
		case KIND_ORCA: return orca_cost(pr.lines, pr.n_lines, a, v);
		case KIND_KARAMOUZAS: return karamouzas_cost(prm, range, a, L, v, pr.ttcp[c]);
		case KIND_RVO: return rvo_cost(prm, range, a, L, v);
		case KIND_MOUSSAID: return moussaid_cost(prm, range, a, L, v);
		case KIND_TTCA_DCA: return ttcadca_cost(prm, range, a, L, v);
		case KIND_PARIS: return paris_cost(prm, range, a, L, v);
		case KIND_VAN_TOLL: return vantoll_cost(prm, range, a, L, v);
		}
		return 0.0f;
	}
}

// CostFunction::GetGradient: analytic for force-based costs (ForceBasedFunction.cpp:32-36)
// and TtcaDca; central differences with delta 0.1 otherwise (costFunction.cpp:36-54).
template <int KMAX, int CLS>
UD V2 gradient_of(const DevPolicy& P, int c, const Self& a, const NbList& L, const Prep<KMAX, CLS>& pr, V2 v) {
	const int kind = P.kind[c];
	if (CLS == CLS_FORCE_GRAD || CLS == CLS_FORCE_SAMPLE || is_force_based(kind)) return (v - pr.target[c]) / a.dt;
	if constexpr (CLS == CLS_GENERIC) {
		if (kind == KIND_TTCA_DCA) return ttcadca_gradient(P.params[c], P.range[c], a, L, v);
		const float d = 0.1f;
		const V2 dx{d, 0.0f}, dy{0.0f, d};
		const float cl = cost_of<KMAX, CLS>(P, c, a, L, pr, v - dx), cr = cost_of<KMAX, CLS>(P, c, a, L, pr, v + dx);
		const float gx = (cr - cl) / (2.0f * d);
		const float cb = cost_of<KMAX, CLS>(P, c, a, L, pr, v - dy), ct = cost_of<KMAX, CLS>(P, c, a, L, pr, v + dy);
		const float gy = (ct - cb) / (2.0f * d);
		return V2{gx, gy};
	}
	return V2{0.0f, 0.0f};
}

// The regular grid of the force-model sampling ranked by sum(coeff_c |v - target_c|^2)
// over the NC costs (the costs without their common factor 0.5 / dt): true, with the
// smallest in v1, when the second smallest lies more than 2^-16 above it (and above
// 2^-60), so that the float costs order the two the same way; false for a near-tie.
template <int NC>
UD bool rank_force_samples(const DevPolicy& P, const V2* target, V2 base, V2 bdir, float dL, int S, int NA, int inc,
                           V2& v1) {
	float coeff[NC];
	V2 tg[NC];
#pragma unroll
	for (int c = 0; c < NC; ++c) {
		coeff[c] = P.coeff[c];
		tg[c] = target[c];
	}
	float s1 = __builtin_inff(), s2 = __builtin_inff();
	auto rank_at = [&](V2 v) {
		float s = 0.0f;
#pragma unroll
		for (int c = 0; c < NC; ++c) s += coeff[c] * sqr_mag(v - tg[c]);
		const bool lt1 = s < s1, lt2 = s < s2;
		s2 = lt1 ? s1 : (lt2 ? s : s2);
		v1 = lt1 ? v : v1;
		s1 = lt1 ? s : s1;
	};
	for (int k = 0; k < NA; ++k) {
		const V2 dir = rotate_cs(bdir, P.ang_cos[k], P.ang_sin[k]);
		if (inc && k == 0) rank_at(base + dir * (dL * 0.0f));
		for (int s = inc ? 2 : 1; s <= S; ++s) rank_at(base + dir * (dL * (float)(inc ? s - 1 : s)));
	}
	return s2 <= kMaxFloat && s2 >= 0x1p-60f && s2 > s1 * (1.0f + 0x1p-16f);
}

// CostFunction::ApproximateGlobalMinimumBySampling (costFunction.cpp:66-164): the first
// strict minimum of sum(coeff * cost) wins; (0,0) if no candidate beats MaxFloat.
template <int KMAX, int CLS>
UD V2 sample_best(const DevPolicy& P, const Self& a, const NbList& L, const Prep<KMAX, CLS>& pr, uint32_t& rng) {
	const V2 base = P.s_base == 1 ? a.vel : V2{0.0f, 0.0f};
	float radius = a.pref_speed;
	if (P.s_radius == 1) radius = a.max_speed;
	else if (P.s_radius == 2) radius = smin(2.0f * a.max_speed, a.max_accel * a.dt);
	V2 bdir{1.0f, 0.0f};
	if (P.s_base_dir == 1) bdir = normalized(a.vel);
	else if (P.s_base_dir == 2) bdir = normalized(a.pref);
	V2 best{0.0f, 0.0f};
	float bestCost = kMaxFloat;
	auto eval = [&](V2 v) {
		float tc = 0.0f;
		for_costs<CLS>(P, [&](int c) { tc += P.coeff[c] * cost_of<KMAX, CLS>(P, c, a, L, pr, v); });
		if (tc < bestCost) { best = v; bestCost = tc; }
	};
	if constexpr (CLS == CLS_FORCE_SAMPLE) {
		if (P.s_type == 0 && P.s_angle_samples <= kMaxAngleSamples) {
			// The same grid in (angle, speed) order, so each direction is rotated once
			// instead of once per speed. The candidate kept is the one the (speed, angle)
			// loop below keeps: the smallest cost, ties to the earlier candidate in that
			// loop's order; a cost not below MaxFloat never wins.
			const int S = P.s_speed_samples, NA = P.s_angle_samples, inc = P.s_include_base;
			const float dL = radius / (float)(inc ? (S - 1) : S);
			// Ranked without the divisions: every cost of the class is a positive
			// multiple of 0.5 |v - target_c|^2 / dt, so sum(coeff_c |v - target_c|^2)
			// orders the candidates as the float cost does wherever two of them differ by
			// more than the rounding of either (each within 2^-20 relative of the exact
			// value, all terms non-negative). When the best and the second best are 2^-16
			// apart or closer (ties, symmetric grids), or the coefficients or dt leave the
			// range where that bound holds, the exact loop below decides.
			bool ranked = P.n_cost >= 1 && a.dt >= 0x1p-20f && a.dt <= 0x1p20f;
			for_costs<CLS>(P, [&](int c) { ranked = ranked && P.coeff[c] >= 0x1p-20f && P.coeff[c] <= 0x1p20f; });
			if (ranked) {
				// one body per cost count, so the loop evaluates only the policy's costs
				V2 v1{0.0f, 0.0f};
				bool decided = false;
				switch (P.n_cost) {
				case 1: decided = rank_force_samples<1>(P, pr.target, base, bdir, dL, S, NA, inc, v1); break;
				case 2: decided = rank_force_samples<2>(P, pr.target, base, bdir, dL, S, NA, inc, v1); break;
				case 3: decided = rank_force_samples<3>(P, pr.target, base, bdir, dL, S, NA, inc, v1); break;
				default: decided = rank_force_samples<4>(P, pr.target, base, bdir, dL, S, NA, inc, v1); break;
				}
				if (decided) {
					float tc = 0.0f;
					for_costs<CLS>(P, [&](int c) { tc += P.coeff[c] * cost_of<KMAX, CLS>(P, c, a, L, pr, v1); });
					return tc < kMaxFloat ? v1 : V2{0.0f, 0.0f};
				}
			}
			int bestIdx = 0x7fffffff;
			auto eval_at = [&](V2 v, int idx) {
				float tc = 0.0f;
				for_costs<CLS>(P, [&](int c) { tc += P.coeff[c] * cost_of<KMAX, CLS>(P, c, a, L, pr, v); });
				if (tc < bestCost || (tc == bestCost && bestIdx != 0x7fffffff && idx < bestIdx)) {
					best = v; bestCost = tc; bestIdx = idx;
				}
			};
			for (int k = 0; k < NA; ++k) {
				const V2 dir = rotate_cs(bdir, P.ang_cos[k], P.ang_sin[k]);
				if (inc && k == 0) eval_at(base + dir * (dL * 0.0f), 0);
				for (int s = inc ? 2 : 1; s <= S; ++s) {
					const float len = dL * (float)(inc ? s - 1 : s);
					eval_at(base + dir * len, inc ? 1 + (s - 2) * NA + k : (s - 1) * NA + k);
				}
			}
			return best;
		}
	}
	if (P.s_type == 1) {
		for (int i = 0; i < P.s_random_samples; ++i) {
			const float ra = rng_uniform(rng, -P.max_angle, P.max_angle);
			const float rl = rng_uniform(rng, 0.0f, radius);
			eval(base + rotate_ccw(bdir, ra) * rl);
		}
	} else {
		const int S = P.s_speed_samples, NA = P.s_angle_samples;
		const float dL = radius / (float)(P.s_include_base ? (S - 1) : S);
		const float start = -P.max_angle;
		const float dA = (P.max_angle - start) / (float)P.s_angle_div;
		for (int s = 1; s <= S; ++s) {
			const float len = dL * (float)(P.s_include_base ? s - 1 : s);
			float ang = start;
			for (int k = 0; k < NA; ++k, ang += dA) {
				const V2 dir = (k < kMaxAngleSamples) ? rotate_cs(bdir, P.ang_cos[k], P.ang_sin[k])
				                                      : rotate_ccw(bdir, ang);
				eval(base + dir * len);
				if (P.s_include_base && s == 1) break;
			}
		}
	}
	return best;
}

// ------------------------------------------------------------------ the step
// The rest of one DoStep for agent k once its neighbour list is known (entry e at
// lst[e * lstride], cnt of them; obstacle edges obs[0 .. n_obs)): the neighbour record,
// Policy::ComputeNewVelocity, contact forces, integration, the outputs and the next
// step's count. BIG: the half-planes of ORCA costs live in caller storage (big_lines,
// big_proj, one per list entry) instead of private arrays of KMAX entries.
template <int KMAX, int CLS, bool BIG>
UD void finish_agent(const StepArgs& A, int k, int4 meta, int pi, int wix, const WorldInfo& wi, const Self& a,
                     const unsigned* lst, int lstride, int cnt, const float4* obs, int n_obs, Line* big_lines,
                     Line* big_proj, int next_key_base) {
	constexpr bool kLong = KMAX == 0;
	constexpr int KL = kLong ? kLongCap : KMAX;
	const State& S = A.in;
	const DevPolicy& P = A.pol[pi];
	if (A.rec_count) {
		A.rec_count[k] = cnt;
		const int m = cnt < A.rec_stride ? cnt : A.rec_stride;
		for (int e = 0; e < m; ++e) {
			const unsigned u = lst[e * lstride];
			const long long o = (long long)k * A.rec_stride + e;
			A.rec_id[o] = S.meta[u & kIndexMask].x;
			A.rec_d2[o] = entry_d2(S, a.pos, wi, u);
			A.rec_seg[o] = (int)(u >> kSegShift);
		}
	}
	if (A.raw_nb) {
		// the adjoint's copy of the list holds raw_stride entries
		const int m = cnt < A.raw_stride ? cnt : A.raw_stride;
		if (cnt > m) atomicOr(A.flags, kFlagRawListTooLong);
		A.raw_count[k] = m;
		for (int e = 0; e < m; ++e) A.raw_nb[(size_t)k * A.raw_stride + e] = lst[e * lstride];
	}
	const NbList L{lst, lstride, cnt, S.pos, S.vel, S.prm, S.pref, wi.W, wi.H, a.pos, obs, n_obs};
	uint32_t rng = (uint32_t)meta.w;

	// Force classes: one pass over the neighbours gathers each neighbour once for the
	// pair forces of every cost (ObjectInteractionForces::ComputeForce,
	// ObjectInteractionForces.cpp:26-48) and for the contact forces (policy.cpp:107-141),
	// each sum still in list order.
	constexpr bool kFused = (CLS == CLS_FORCE_GRAD || CLS == CLS_FORCE_SAMPLE);
	V2 fsum[kFused ? 4 : 1];
	V2 contact{0.0f, 0.0f};
	if constexpr (kFused) {
		float rsq[4];
#pragma unroll
		for (int c = 0; c < 4; ++c) {
			fsum[c] = V2{0.0f, 0.0f};
			rsq[c] = P.range[c] * P.range[c];
		}
		const bool do_contact = P.contact_scale > 0.0f;
		for (int e = 0; e < cnt; ++e) {
			const Nb n = L.get(e);
#pragma unroll
			for (int c = 0; c < 4; ++c) {
				if (c < P.n_cost && P.kind[c] != KIND_GOAL_REACHING && n.dsq <= rsq[c]) {
					const V2 f = (P.kind[c] == KIND_POWER_LAW)
					                 ? power_law_pair(P.params[c], a, n)
					                 : social_force_pair(P.params[c], a, n, L.pref_of(n.j), P.view_cos);
					fsum[c] = fsum[c] + f;
				}
			}
			if (do_contact) {
				const V2 diff = a.pos - n.pos;
				const float inter = a.radius + n.radius - mag(diff);
				if (inter > 0.0f) contact = contact + normalized(diff) * (P.contact_scale * inter);
			}
		}
	}

	// Policy::ComputeNewVelocity (policy.cpp:43-76)
	V2 next{0.0f, 0.0f};
	if (!is_zero(a.pref)) {
		Prep<KMAX, CLS> pr;
		pr.n_lines = 0;
		if constexpr (CLS == CLS_ORCA) {
			Line proj_priv[BIG ? 1 : KL];
			const Lines proj{BIG ? big_proj : proj_priv, 1};
			V2 best;
			if constexpr (!kLong && KMAX <= 16) {
				// half-planes in registers; only the rare projected lines of
				// linearProgram3 go to private memory
				RegLines<KMAX> R;
				orca_lines_reg<KMAX>(P.params[0], P.range[0], a, L, R);
				best = orca_solve_reg<KMAX>(R, a, proj);
			} else {
				Line lines_priv[BIG ? 1 : KL];
				pr.lines = Lines{BIG ? big_lines : lines_priv, 1};
				pr.n_lines = orca_lines(P.params[0], P.range[0], a, L, pr.lines);
				best = orca_solve(pr.lines, pr.n_lines, a, proj);
			}
			const V2 acc = clamp_vec((best - a.vel) / smax(a.relaxation, a.dt), a.max_accel);
			next = clamp_vec(a.vel + acc * a.dt, a.max_speed);
		} else {
			Line lines_priv[CLS == CLS_GENERIC && !BIG ? KL : 1];
			Line proj_priv[CLS == CLS_GENERIC && !BIG ? KL : 1];
			pr.lines = Lines{BIG ? big_lines : lines_priv, 1};
			for_costs<CLS>(P, [&](int c) {
				const int kind = P.kind[c];
				if constexpr (kFused) {
					// ForceBasedFunction::ComputeTargetVelocity (ForceBasedFunction.cpp:49-52)
					const V2 f = (kind == KIND_GOAL_REACHING) ? (a.pref - a.vel) / smax(a.relaxation, a.dt)
					                                          : fsum[c] + V2{0.0f, 0.0f};
					pr.target[c] = a.vel + f / P.params[c][0] * a.dt;
				} else if (CLS != CLS_GENERIC || is_force_based(kind)) pr.target[c] = target_velocity(P, c, a, L);
				else if (kind == KIND_ORCA) pr.n_lines = orca_lines(P.params[c], P.range[c], a, L, pr.lines);
				else if (kind == KIND_KARAMOUZAS)
					pr.ttcp[c] = time_to_first_collision(L, a.pos, a.pref, a.radius, P.range[c], true);
			});
			V2 acc;
			if (CLS == CLS_FORCE_GRAD || (CLS == CLS_GENERIC && P.mode == MODE_GRADIENT)) {
				if (CLS == CLS_FORCE_GRAD && P.mode == MODE_GLOBAL_FORCE) {
					acc = (pr.target[0] - a.vel) / smax(a.relaxation, a.dt);
				} else {
					V2 tg{0.0f, 0.0f};
					for_costs<CLS>(P, [&](int c) { tg = tg + P.coeff[c] * gradient_of<KMAX, CLS>(P, c, a, L, pr, a.vel); });
					acc = -1.0f * tg;
					if (P.grad_iters > 1) {
						// the descent extension (umans_policy::gradient_iterations): the
						// first iteration's gradient is the one just summed
						const float eta = P.grad_step * a.dt;
						V2 v = a.vel - tg * eta;
						for (int it = 1; it < P.grad_iters; ++it) {
							V2 g{0.0f, 0.0f};
							for_costs<CLS>(P, [&](int c) { g = g + P.coeff[c] * gradient_of<KMAX, CLS>(P, c, a, L, pr, v); });
							v = v - g * eta;
						}
						acc = (v - a.vel) / a.dt;
					}
				}
			} else {
				V2 best;
				if (CLS == CLS_GENERIC && P.mode == MODE_GLOBAL_ORCA)
					best = orca_solve(pr.lines, pr.n_lines, a, Lines{BIG ? big_proj : proj_priv, 1});
				else if (CLS == CLS_GENERIC && P.mode == MODE_GLOBAL_FORCE) best = pr.target[0];
				else best = sample_best<KMAX, CLS>(P, a, L, pr, rng);
				acc = (best - a.vel) / smax(a.relaxation, a.dt);
			}
			acc = clamp_vec(acc, a.max_accel);
			next = clamp_vec(a.vel + acc * a.dt, a.max_speed);
		}
	}

	// Policy::ComputeContactForces (policy.cpp:107-141)
	if (!kFused && P.contact_scale > 0.0f) {
		for (int e = 0; e < cnt; ++e) {
			const Nb n = L.get(e);
			const V2 diff = a.pos - n.pos;
			const float inter = a.radius + n.radius - mag(diff);
			if (inter > 0.0f) contact = contact + normalized(diff) * (P.contact_scale * inter);
		}
		for (int e = 0; e < n_obs; ++e) {
			const V2 diff = a.pos - nearest_point(a.pos, seg_a(obs[e]), seg_b(obs[e]), true);
			const float inter = a.radius - mag(diff);
			if (inter > 0.0f) contact = contact + normalized(diff) * (P.contact_scale * inter);
		}
	}

	// Agent::UpdateVelocityAndPosition (agent.cpp:71-93) + toric wrap (worldToric.cpp:81-108)
	const V2 vel = next + ((contact / 80.0f) * a.dt);
	V2 view = v2(S.viewdir[k]);
	const V2 vd = (2.0f * vel + a.pref) / 3.0f;
	if ((double)sqr_mag(vd) > 0.01) view = normalized(vd);
	V2 pos = a.pos + (vel * a.dt);
	if (wi.world_type == 1) {
		const float hw = 0.5f * wi.W, hh = 0.5f * wi.H;
		if (pos.x > hw) pos.x -= wi.W; else if (pos.x < -hw) pos.x += wi.W;
		if (pos.y > hh) pos.y -= wi.H; else if (pos.y < -hh) pos.y += wi.H;
	}
	A.out.pos[k] = f2(pos);
	A.out.vel[k] = f2(vel);
	A.out.viewdir[k] = f2(view);
	if (A.out.meta) {
		// a full copy of the record
		A.out.goal[k] = S.goal[k];
		A.out.prm[k] = S.prm[k];
		A.out.pref[k] = f2(a.pref);
		meta.w = (int)rng;
		A.out.meta[k] = meta;
	} else if (rng != (uint32_t)meta.w) {
		// the static fields stay where they were read (the host swaps the buffers): only
		// the sampling engine's state changes, in place (no other agent reads it)
		reinterpret_cast<int*>(S.meta + k)[3] = (int)rng;
	}
	if (A.next_count) count_next(A, wi, wix, k, f2(pos), next_key_base);
}

// One lane per agent of the class's list entries [A.base, A.base + A.n).
template <int KMAX, int BS, int CLS>
__global__ __launch_bounds__(BS) void k_step(StepArgs A) {
	// KMAX = 0: long lists, collected in global buffers and ranked by the wave (BS = 64)
	constexpr bool kLong = KMAX == 0;
	constexpr int KL = kLong ? kLongCap : KMAX;    // the longest list a lane can hold
	static_assert(!kLong || BS == 64, "long lists are ranked per wave: one wave per block");
	constexpr bool kShort = !kLong && KMAX <= kShortList;
	constexpr int KC = short_buffer<KMAX, kStepMinKeys>();
	__shared__ float s_d2[kLong || kShort ? 1 : KMAX * BS];
	__shared__ unsigned s_nb[kLong || kShort ? 1 : KMAX * BS];
	// short lists: each wave's region of KC keys per lane, its list written over the start
	__shared__ unsigned long long s_buf[kShort ? KC * BS : 1];
	// The ORCA half-planes live in private memory: kept in LDS they bound the class to
	// ~2 waves per SIMD, and the private copy (L1-resident) measured faster at 6.
	const int tid = threadIdx.x;
	// long lists: apw agents per wave, lanes tid >= apw hold none
	const int apw = kLong && A.long_apw > 0 ? A.long_apw : BS;
	const int t = xcd_tile(blockIdx.x, gridDim.x) * apw + tid;
	const int n_here = A.agent_count ? (int)*A.agent_count : A.n;
	const int base = A.dev_base ? (int)*A.dev_base : A.base;
	const bool active = tid < apw && t < n_here;
	// a wave without agents (most of a launch over a list counted on the device) leaves
	// at once: it takes part in no wave-level step and adds nothing to the stats
	if (__ballot(active) == 0ull) return;
	const int k = active ? (A.agent_list ? A.agent_list[base + t] : A.class_list[base + t]) : 0;
	int total = 0, cnt = 0;
	const State& S = A.in;
	int4 meta = make_int4(0, 0, 0, 0);
	int pi = 0, wix = 0;
	Self a;
	WorldInfo wi{};
	if (active) {
		// only what the search needs is loaded before it (fewer registers live across it)
		const int my = S.meta[k].y;
		pi = my & kPolicyMask;
		wix = my >> kPolicyBits;
		wi = A.worlds[wix];
		const GridInfo g = A.grids[wix];
		a.pos = v2(S.pos[k]);
		if constexpr (kLong) {
			total = collect_neighbors(A, wi, g, k, a.pos, A.pol[pi].query_range, base + t);
		} else if constexpr (kShort) {
			unsigned long long* region = s_buf + (tid >> 6) * (KC * 64);
			cnt = find_short<KMAX, kStepMinKeys>(A, wi, g, k, a.pos, A.pol[pi].query_range, region + (tid & 63),
			                       reinterpret_cast<unsigned*>(region) + (tid & 63), total);
		} else {
			cnt = find_neighbors<KMAX, BS>(A, wi, g, k, a.pos, A.pol[pi].query_range, s_d2 + tid, s_nb + tid, total);
		}
		meta = S.meta[k];
		const float4 prm = S.prm[k];
		a.vel = v2(S.vel[k]);
		a.pref = v2(S.pref[k]);
		a.radius = prm.x; a.pref_speed = prm.y; a.max_speed = prm.z; a.max_accel = prm.w;
		a.dt = wi.dt;
		a.relaxation = A.pol[pi].relaxation;
	}
	// long lists: a lane with more than kLongCap in range is handed on; the others are
	// ranked by the wave (every lane takes part, active or not)
	const int kcap_eff = A.kcap < KL ? A.kcap : KL;
	if constexpr (kLong) {
		const bool keep = active && total <= kLongCap;
		rank_long_lists(A, base + t, keep ? total : 0, kcap_eff, a.pos, wi);
		cnt = keep ? (total < kcap_eff ? total : kcap_eff) : 0;
	}
	// obstacle edges: only the generic class runs when a world has obstacles
	float4 obs_priv[CLS == CLS_GENERIC ? kMaxObstacleEdges : 1];
	int n_obs = 0;
	bool many_edges = false;
	if constexpr (CLS == CLS_GENERIC) {
		if (active && A.n_edges > 0) {
			int obs_total = 0;
			n_obs = find_obstacles(A, wi, a.pos, A.pol[pi].query_range, obs_priv, obs_total);
			many_edges = obs_total > n_obs;
			if (many_edges && !A.huge_list) atomicAdd(&A.stats[stat_copy() * kStatLine + 2], 1ull);
		}
	}
	// a list that outgrew this launch's capacity is redone by the next launch (with a cap
	// or below 128 entries) or, with no cap, by the long-list launch k_step_huge, as is
	// an agent with more obstacle edges in range than a lane keeps
	const bool beyond = kLong ? total > kLongCap : (A.tier_final ? total > KMAX : false);
	const bool to_huge = active && A.huge_list && ((A.uncapped && beyond) || many_edges);
	const bool handover = !to_huge && (kLong ? total > kLongCap : (!A.tier_final && total > KMAX));
	// stats: one atomic per wave
	{
		const bool final_here = !to_huge && (kLong ? !handover : A.tier_final != 0);
		const unsigned long long ov = __ballot(active && final_here && total > A.kcap);
		const int mx = wave_max_i(total);
		if ((tid & 63) == 0) {
			const int c = __popcll(ov);
			unsigned long long* sc = A.stats + stat_copy() * kStatLine;
			if (c) atomicAdd(&sc[0], (unsigned long long)c);
			atomicMax(&sc[1], (unsigned long long)mx);
		}
	}
	if (!active) return;
	if (to_huge) {
		A.huge_list[atomicAdd(A.huge_count, 1u)] = k;
		return;
	}
	if (handover) {
		A.ovf_list[A.ovf_base + atomicAdd(A.ovf_count, 1u)] = k;
		return;
	}
	// the lane's sorted list: its LDS column, or its row of the long-list buffer
	const unsigned* lst = kLong ? A.long_u + (base + t)
	                            : (kShort ? reinterpret_cast<const unsigned*>(s_buf + (tid >> 6) * (KC * 64)) + (tid & 63)
	                                      : s_nb + tid);
	const int lstride = kLong ? A.long_stride : (kShort ? 64 : BS);
	finish_agent<KMAX, CLS, false>(A, k, meta, pi, wix, wi, a, lst, lstride, cnt, obs_priv, n_obs, nullptr, nullptr,
	                               A.next_key_base);
}


// ----------------------------------------------------------- very long lists
// The agents k_step hands on: with no cap, those with more than 128 neighbours in range
// (the reference keeps every neighbour, AgentKDTree.cpp:35-59), and in any world those
// with more than kMaxObstacleEdges obstacle edges in range (worldBase.cpp:62-77). One
// wave per agent, in a grid-stride loop over the handed-on agents (launched after every
// class kernel of the step, so the count is final): the lanes collect the candidates in
// range into LDS as packed keys (segment | float(d2) | index), sort them with a bitonic
// network over the next power of two, order the runs of equal floats by their exact
// doubles and agent IDs, and one lane runs the rest of the step with the generic code
// path, the half-planes of ORCA costs and the obstacle edges in the block's slice of
// scratch. Lists longer than kHugeMax are not kept whole: kFlagListTooLong tells the host.
struct HugeArgs {
	StepArgs s;
	const int* cls_slot;     // sort-key slot of each policy class (the next step's keys)
	int slot_cells;
	float4* edges;           // edge_cap per block
	int edge_cap;
	Line* lines;             // 2 * kHugeMax per block: half-planes, projected lines
};

#ifndef UMANS_STEP_KERNELS_ONLY
__global__ __launch_bounds__(64) void k_step_huge(HugeArgs H) {
	__shared__ unsigned long long s_key[kHugeMax];
	__shared__ unsigned s_u[kHugeMax];
	__shared__ unsigned s_cnt;
	const StepArgs& A = H.s;
	const State& S = A.in;
	const int lane = threadIdx.x;
	const unsigned n_huge = *A.huge_count;
	float4* edges = H.edges + (size_t)blockIdx.x * H.edge_cap;
	Line* lines = H.lines + (size_t)blockIdx.x * 2 * kHugeMax;
	for (unsigned h = blockIdx.x; h < n_huge; h += gridDim.x) {
		const int k = A.huge_list[h];
		const int4 meta = S.meta[k];
		const int pi = meta.y & kPolicyMask, wix = meta.y >> kPolicyBits;
		const WorldInfo wi = A.worlds[wix];
		const GridInfo g = A.grids[wix];
		const DevPolicy& P = A.pol[pi];
		const float4 prm = S.prm[k];
		Self a;
		a.pos = v2(S.pos[k]);
		a.vel = v2(S.vel[k]);
		a.pref = v2(S.pref[k]);
		a.radius = prm.x; a.pref_speed = prm.y; a.max_speed = prm.z; a.max_accel = prm.w;
		a.dt = wi.dt;
		a.relaxation = P.relaxation;
		const float range = P.query_range;
		// the queries of find_neighbors, the candidates of each row range spread over the lanes
		if (lane == 0) s_cnt = 0u;
		__syncthreads();
		const double r2 = (double)(range * range);
		const float rs = range * 1.0001f + 1e-3f;
		for (int s = 0; s < 5; ++s) {
			float dx = 0.0f, dy = 0.0f;
			if (s > 0) {
				if (wi.world_type != 1) break;
				if (s == 1) { if (!((double)(a.pos.x - range) < -0.5 * (double)wi.W)) continue; dx = wi.W; }
				else if (s == 2) { if (!((double)(a.pos.x + range) > 0.5 * (double)wi.W)) continue; dx = -wi.W; }
				else if (s == 3) { if (!((double)(a.pos.y - range) < -0.5 * (double)wi.H)) continue; dy = wi.H; }
				else { if (!((double)(a.pos.y + range) > 0.5 * (double)wi.H)) continue; dy = -wi.H; }
			}
			const V2 q{a.pos.x + dx, a.pos.y + dy};
			const float m = rs + (fabsf(q.x) + fabsf(q.y)) * 1e-6f;
			const int cx0 = bin_axis(q.x - m, g.ox, g.invx, g.ncx), cx1 = bin_axis(q.x + m, g.ox, g.invx, g.ncx);
			const int cy0 = bin_axis(q.y - m, g.oy, g.invy, g.ncy), cy1 = bin_axis(q.y + m, g.oy, g.invy, g.ncy);
			const double qx = (double)q.x, qy = (double)q.y;
			for (int cy = cy0; cy <= cy1; ++cy) {
				const int crow = wi.cell_offset + cy * g.ncx;
				const unsigned b = A.cell_start[crow + cx0], e = A.cell_start[crow + cx1 + 1];
				for (unsigned j = b + lane; j < e; j += 64) {
					const float2 pj = S.pos[j];
					const double ddx = qx - (double)pj.x;
					const double ddy = qy - (double)pj.y;
					const double d2 = ddx * ddx + ddy * ddy;
					if (!(d2 < r2) || (int)j == k) continue;
					const unsigned slot = atomicAdd(&s_cnt, 1u);
					if (slot < (unsigned)kHugeMax) s_key[slot] = long_key(s, d2, j);
				}
			}
		}
		__syncthreads();
		const int total = (int)s_cnt;
		const int n = total < kHugeMax ? total : kHugeMax;
		int P2 = 64;
		while (P2 < n) P2 <<= 1;
		for (int i = n + lane; i < P2; i += 64) s_key[i] = kKeyPad;
		__syncthreads();
		// bitonic sort of the packed keys, ascending
		for (int w = 2; w <= P2; w <<= 1) {
			for (int j = w >> 1; j > 0; j >>= 1) {
				for (int i = lane; i < P2; i += 64) {
					const int ixj = i ^ j;
					if (ixj > i) {
						const unsigned long long x = s_key[i], y = s_key[ixj];
						if ((x > y) == ((i & w) == 0)) { s_key[i] = y; s_key[ixj] = x; }
					}
				}
				__syncthreads();
			}
		}
		for (int i = lane; i < n; i += 64) s_u[i] = key_entry(s_key[i]);
		__syncthreads();
		if (lane == 0) {
			// runs of equal segment and float(d2) in the canonical order (insertion by exact
			// d2 and ID; the runs are a few entries long)
			for (int i = 0; i + 1 < n;) {
				int r = i + 1;
				while (r < n && same_float_key(s_key[i], s_key[r])) ++r;
				for (int x = i + 1; x < r; ++x) {
					const unsigned u = s_u[x];
					int y = x;
					while (y > i && entry_before(S, a.pos, wi, u, s_u[y - 1])) {
						s_u[y] = s_u[y - 1];
						--y;
					}
					s_u[y] = u;
				}
				i = r;
			}
			int n_obs = 0;
			if (A.n_edges > 0) {
				int obs_total = 0;
				n_obs = find_obstacles(A, wi, a.pos, range, edges, obs_total, H.edge_cap);
			}
			const bool too_long = A.uncapped ? total > kHugeMax : total > A.kcap;
			unsigned long long* sc = A.stats + stat_copy() * kStatLine;
			if (too_long) atomicAdd(&sc[0], 1ull);
			atomicMax(&sc[1], (unsigned long long)total);
			// the LDS list holds kHugeMax of the candidates, in the order they were found
			if (total > kHugeMax) atomicOr(A.flags, kFlagListTooLong);
			const int cnt = A.uncapped ? n : (n < A.kcap ? n : A.kcap);
			// the next step's key: the slot of the agent's class (with obstacles every
			// agent is in the generic class)
			const int cls = A.n_edges > 0 ? (int)CLS_GENERIC : P.cls;
			finish_agent<0, CLS_GENERIC, true>(A, k, meta, pi, wix, wi, a, s_u, 1, cnt, edges, n_obs, lines,
			                                   lines + kHugeMax, H.cls_slot[cls] * H.slot_cells);
		}
		__syncthreads();          // the next agent reuses the LDS lists
	}
}
#endif


// --------------------------------------------------------- cooperative sampling
// CostFunction::ApproximateGlobalMinimumBySampling (costFunction.cpp:66-164) for the
// regular grid with the candidates of the 64 agents of a wave spread over its 64 lanes:
// candidate c of agent i is work item items[i] + c. Per item, the agent's state and
// its neighbour list are read from LDS and the neighbours' state from memory; lanes
// working on one agent read the same neighbours, so those loads are broadcasts instead
// of 64 separate gathers. The winner is the first strict minimum in candidate order,
// as the sequential loop finds it: the smallest key (cost in an order-preserving
// encoding, candidate index) wins, costs not below MaxFloat never do.
UD unsigned long long cost_key(float tc, int c) {
	if (!(tc < kMaxFloat)) return ~0ull;
	unsigned u = __float_as_uint(tc == 0.0f ? 0.0f : tc);   // -0 and +0 are equal costs
	u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
	return ((unsigned long long)u << 32) | (unsigned)c;
}

// Candidate c of the regular grid (the loop order of sample_best).
UD V2 regular_candidate(const DevPolicy& P, int c, V2 base, V2 bdir, float dL) {
	int s, k;
	if (P.s_include_base) {
		if (c == 0) { s = 1; k = 0; }
		else { s = 2 + (c - 1) / P.s_angle_samples; k = (c - 1) % P.s_angle_samples; }
	} else {
		s = 1 + c / P.s_angle_samples;
		k = c % P.s_angle_samples;
	}
	const float len = dL * (float)(P.s_include_base ? s - 1 : s);
	return base + rotate_cs(bdir, P.ang_cos[k], P.ang_sin[k]) * len;
}

UD int regular_count(const DevPolicy& P) {
	return P.s_include_base ? 1 + (P.s_speed_samples - 1) * P.s_angle_samples : P.s_speed_samples * P.s_angle_samples;
}

// The cost of candidate v for cost q of policy Q: KIND >= 0 when every policy of the
// class has that one cost (only its code is compiled in: 127 instead of 181 registers for
// Karamouzas, twice the resident waves), -1 for any mix.
template <int KMAX, int KIND>
UD float coop_cost(const DevPolicy& Q, int q, const Self& b, const NbList& L, const Prep<KMAX, CLS_GENERIC>& pr, V2 v) {
	if constexpr (KIND == KIND_KARAMOUZAS) return karamouzas_cost(Q.params[q], Q.range[q], b, L, v, pr.ttcp[q]);
	else if constexpr (KIND == KIND_RVO) return rvo_cost(Q.params[q], Q.range[q], b, L, v);
	else if constexpr (KIND == KIND_MOUSSAID) return moussaid_cost(Q.params[q], Q.range[q], b, L, v);
	else if constexpr (KIND == KIND_TTCA_DCA) return ttcadca_cost(Q.params[q], Q.range[q], b, L, v);
	else if constexpr (KIND == KIND_PARIS) return paris_cost(Q.params[q], Q.range[q], b, L, v);
	else if constexpr (KIND == KIND_VAN_TOLL) return vantoll_cost(Q.params[q], Q.range[q], b, L, v);
	else return cost_of<KMAX, CLS_GENERIC>(Q, q, b, L, pr, v);
}

// Resident waves per SIMD asked of the compiler: the Karamouzas instance with 16 agents
// per wave (113 VGPRs, 4 waves) measured 5 % faster held to 72 VGPRs and 7 waves (140
// bytes of spills; 6 and 8 waves gained less or lost). With 32 or 64 agents per wave
// their LDS caps the waves below that anyway; the other instances keep their allocation.
constexpr int coop_waves_per_simd(int kind, int apw) { return kind == KIND_KARAMOUZAS && apw == 16 ? 7 : 1; }

// APW agents per wave (64 or 32): fewer agents per wave make more waves, for crowds too
// small to fill the SIMDs at 64 (the candidates of the wave's agents still cover all lanes).
template <int KMAX, int APW, int KIND>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(coop_waves_per_simd(KIND, APW)))) void k_step_coop(StepArgs A) {
	static_assert(APW == 64 || APW == 32 || APW == 16, "agents per wave");
	constexpr bool kShort = KMAX <= kShortList;
	constexpr int KC = short_buffer<KMAX, kCoopMinKeys>();
	// per-agent arrays hold the wave's APW agents (lanes tid < APW): with 32 agents per
	// wave half the LDS of 64, so twice the waves fit a CU
	__shared__ float s_d2[kShort ? 1 : KMAX * APW];
	__shared__ unsigned s_nbl[kShort ? 1 : KMAX * APW];
	__shared__ unsigned long long s_buf[kShort ? KC * APW : 1];
	// the lists, agent-interleaved (stride APW): short lists are written over the keys
	unsigned* const s_nb = kShort ? reinterpret_cast<unsigned*>(s_buf) : s_nbl;
	__shared__ float s_self[12][APW];    // pos, vel, pref, radius, pref speed, max speed, max accel, dt, relaxation
	__shared__ float s_ttcp[4][APW];
	__shared__ float s_wh[2][APW];
	__shared__ int s_cnt[APW], s_pi[APW], s_items[APW + 1];
	__shared__ unsigned long long s_key[APW];
	const int tid = threadIdx.x;
	const int t = xcd_tile(blockIdx.x, gridDim.x) * APW + tid;
	const int n_here = A.agent_count ? (int)*A.agent_count : A.n;
	const bool active = tid < APW && t < n_here;
	// one wave per block: a block without agents (a launch sized for a count known only
	// on the device) leaves at once, before any barrier
	if (__ballot(active) == 0ull) return;
	const int k = active ? A.class_list[(A.dev_base ? (int)*A.dev_base : A.base) + t] : 0;
	int total = 0, cnt = 0;
	const State& S = A.in;
	int4 meta = make_int4(0, 0, 0, 0);
	int pi = 0;
	Self a{};
	WorldInfo wi{};
	if (active) {
		meta = S.meta[k];
		pi = meta.y & kPolicyMask;
		const int wix = meta.y >> kPolicyBits;
		wi = A.worlds[wix];
		const GridInfo g = A.grids[wix];
		const float4 prm = S.prm[k];
		a.pos = v2(S.pos[k]);
		a.vel = v2(S.vel[k]);
		a.pref = v2(S.pref[k]);
		a.radius = prm.x; a.pref_speed = prm.y; a.max_speed = prm.z; a.max_accel = prm.w;
		a.dt = wi.dt;
		a.relaxation = A.pol[pi].relaxation;
		if constexpr (kShort)
			cnt = find_short<KMAX, kCoopMinKeys, APW>(A, wi, g, k, a.pos, A.pol[pi].query_range, s_buf + tid, s_nb + tid, total);
		else
			cnt = find_neighbors<KMAX, APW>(A, wi, g, k, a.pos, A.pol[pi].query_range, s_d2 + tid, s_nb + tid, total);
	}
	// with no cap, a list longer than KMAX is redone whole by k_step_huge; the lane still
	// takes part in the wave's barriers but has no candidates
	const bool to_huge = active && A.huge_list && A.uncapped && total > KMAX;
	{
		const unsigned long long ov = __ballot(active && !to_huge && total > A.kcap);
		const int mx = wave_max_i(total);
		if (tid == 0) {
			const int c = __popcll(ov);
			unsigned long long* sc = A.stats + stat_copy() * kStatLine;
			if (c) atomicAdd(&sc[0], (unsigned long long)c);
			atomicMax(&sc[1], (unsigned long long)mx);
		}
	}
	const DevPolicy& P = A.pol[pi];
	if (active && !to_huge && A.rec_count) {
		A.rec_count[k] = cnt;
		const int m = cnt < A.rec_stride ? cnt : A.rec_stride;
		for (int e = 0; e < m; ++e) {
			const unsigned u = s_nb[e * APW + tid];
			const long long o = (long long)k * A.rec_stride + e;
			A.rec_id[o] = S.meta[u & kIndexMask].x;
			A.rec_d2[o] = entry_d2(S, a.pos, wi, u);
			A.rec_seg[o] = (int)(u >> kSegShift);
		}
	}
	// (lanes tid >= APW hold no agent and never read their list)
	const NbList L{s_nb + (tid < APW ? tid : 0), APW, cnt, S.pos, S.vel, S.prm, S.pref, wi.W, wi.H, a.pos, nullptr, 0};
	// Policy::ComputeNewVelocity (policy.cpp:43-76): agents with a zero preferred
	// velocity keep a zero new velocity and have no candidates
	const bool sample = active && !to_huge && !is_zero(a.pref);
	const int ns = sample ? regular_count(P) : 0;
	float ttcp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	if (sample) {
		for (int c = 0; c < P.n_cost; ++c)
			if (P.kind[c] == KIND_KARAMOUZAS) ttcp[c] = time_to_first_collision(L, a.pos, a.pref, a.radius, P.range[c], true);
	}
	const float self_v[12] = {a.pos.x, a.pos.y, a.vel.x, a.vel.y, a.pref.x, a.pref.y, a.radius, a.pref_speed,
	                          a.max_speed, a.max_accel, a.dt, a.relaxation};
	if (tid < APW) {
#pragma unroll
		for (int q = 0; q < 12; ++q) s_self[q][tid] = self_v[q];
#pragma unroll
		for (int q = 0; q < 4; ++q) s_ttcp[q][tid] = ttcp[q];
		s_wh[0][tid] = wi.W;
		s_wh[1][tid] = wi.H;
		s_cnt[tid] = cnt;
		s_pi[tid] = pi;
		s_key[tid] = ~0ull;
	}
	// exclusive prefix sum of the candidate counts over the wave (lanes tid >= APW have none)
	int incl = ns;
	for (int o = 1; o < 64; o <<= 1) {
		const int y = __shfl_up(incl, o);
		if (tid >= o) incl += y;
	}
	if (tid < APW) s_items[tid + 1] = incl;
	if (tid == 0) s_items[0] = 0;
	__syncthreads();
	const int n_items = s_items[APW];

	int ag = -1;                                   // agent whose state the lane holds
	unsigned long long local = ~0ull;
	Self b{};
	int bcnt = 0, bpi = 0;
	float bW = 0.0f, bH = 0.0f;
	Prep<KMAX, CLS_GENERIC> pr;
	pr.n_lines = 0;
	V2 base{0.0f, 0.0f}, bdir{1.0f, 0.0f};
	float dL = 0.0f;
	int cur = 0;
	for (int it = tid; it < n_items; it += 64) {
		while (it >= s_items[cur + 1]) ++cur;
		if (cur != ag) {
			if (ag >= 0 && local != ~0ull) atomicMin(&s_key[ag], local);
			local = ~0ull;
			ag = cur;
			b.pos = V2{s_self[0][ag], s_self[1][ag]};
			b.vel = V2{s_self[2][ag], s_self[3][ag]};
			b.pref = V2{s_self[4][ag], s_self[5][ag]};
			b.radius = s_self[6][ag]; b.pref_speed = s_self[7][ag]; b.max_speed = s_self[8][ag];
			b.max_accel = s_self[9][ag]; b.dt = s_self[10][ag]; b.relaxation = s_self[11][ag];
			bcnt = s_cnt[ag];
			bpi = s_pi[ag];
			bW = s_wh[0][ag];
			bH = s_wh[1][ag];
#pragma unroll
			for (int q = 0; q < 4; ++q) pr.ttcp[q] = s_ttcp[q][ag];
			const DevPolicy& Q = A.pol[bpi];
			// sample_best's set-up: base, radius, base direction, speed step
			base = Q.s_base == 1 ? b.vel : V2{0.0f, 0.0f};
			float radius = b.pref_speed;
			if (Q.s_radius == 1) radius = b.max_speed;
			else if (Q.s_radius == 2) radius = smin(2.0f * b.max_speed, b.max_accel * b.dt);
			bdir = V2{1.0f, 0.0f};
			if (Q.s_base_dir == 1) bdir = normalized(b.vel);
			else if (Q.s_base_dir == 2) bdir = normalized(b.pref);
			dL = radius / (float)(Q.s_include_base ? (Q.s_speed_samples - 1) : Q.s_speed_samples);
		}
		const DevPolicy& Q = A.pol[bpi];
		const int c = it - s_items[ag];
		const V2 v = regular_candidate(Q, c, base, bdir, dL);
		const NbList Lb{s_nb + ag, APW, bcnt, S.pos, S.vel, S.prm, S.pref, bW, bH, b.pos, nullptr, 0};
		float tc = 0.0f;
		if constexpr (KIND >= 0) {
			// the class's policies all have this one cost: index 0, so its per-agent terms
			// stay in registers
			tc += Q.coeff[0] * coop_cost<KMAX, KIND>(Q, 0, b, Lb, pr, v);
		} else {
			for (int q = 0; q < Q.n_cost; ++q) tc += Q.coeff[q] * coop_cost<KMAX, KIND>(Q, q, b, Lb, pr, v);
		}
		const unsigned long long key = cost_key(tc, c);
		if (key < local) local = key;
	}
	if (ag >= 0 && local != ~0ull) atomicMin(&s_key[ag], local);
	__syncthreads();
	if (!active) return;
	if (to_huge) {
		A.huge_list[atomicAdd(A.huge_count, 1u)] = k;
		return;
	}

	V2 next{0.0f, 0.0f};
	if (sample) {
		const unsigned long long key = s_key[tid];
		V2 best{0.0f, 0.0f};
		if (key != ~0ull) {
			const V2 bs = P.s_base == 1 ? a.vel : V2{0.0f, 0.0f};
			float radius = a.pref_speed;
			if (P.s_radius == 1) radius = a.max_speed;
			else if (P.s_radius == 2) radius = smin(2.0f * a.max_speed, a.max_accel * a.dt);
			V2 bd{1.0f, 0.0f};
			if (P.s_base_dir == 1) bd = normalized(a.vel);
			else if (P.s_base_dir == 2) bd = normalized(a.pref);
			const float dl = radius / (float)(P.s_include_base ? (P.s_speed_samples - 1) : P.s_speed_samples);
			best = regular_candidate(P, (int)(key & 0xffffffffu), bs, bd, dl);
		}
		const V2 acc = clamp_vec((best - a.vel) / smax(a.relaxation, a.dt), a.max_accel);
		next = clamp_vec(a.vel + acc * a.dt, a.max_speed);
	}
	// Policy::ComputeContactForces (policy.cpp:107-141)
	V2 contact{0.0f, 0.0f};
	if (P.contact_scale > 0.0f) {
		for (int e = 0; e < cnt; ++e) {
			const Nb n = L.get(e);
			const V2 diff = a.pos - n.pos;
			const float inter = a.radius + n.radius - mag(diff);
			if (inter > 0.0f) contact = contact + normalized(diff) * (P.contact_scale * inter);
		}
	}
	// Agent::UpdateVelocityAndPosition (agent.cpp:71-93) + toric wrap (worldToric.cpp:81-108)
	const V2 vel = next + ((contact / 80.0f) * a.dt);
	V2 view = v2(S.viewdir[k]);
	const V2 vd = (2.0f * vel + a.pref) / 3.0f;
	if ((double)sqr_mag(vd) > 0.01) view = normalized(vd);
	V2 pos = a.pos + (vel * a.dt);
	if (wi.world_type == 1) {
		const float hw = 0.5f * wi.W, hh = 0.5f * wi.H;
		if (pos.x > hw) pos.x -= wi.W; else if (pos.x < -hw) pos.x += wi.W;
		if (pos.y > hh) pos.y -= wi.H; else if (pos.y < -hh) pos.y += wi.H;
	}
	A.out.pos[k] = f2(pos);
	A.out.vel[k] = f2(vel);
	A.out.viewdir[k] = f2(view);
	if (A.out.meta) {             // a full copy (else the static fields stay where they are)
		A.out.goal[k] = S.goal[k];
		A.out.prm[k] = S.prm[k];
		A.out.pref[k] = f2(a.pref);
		A.out.meta[k] = meta;
	}
	if (A.next_count) count_next(A, wi, meta.y >> kPolicyBits, k, f2(pos), A.next_key_base);
}

} // namespace umans
