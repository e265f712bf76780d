// sparse_build.hip — building resident sparse scenes on the GPU.
//
// rkfm_sparse_build voxelises analytic primitives the way voxelize_sdf does
// (crates/rkf-core/src/voxelize_object.rs:53-182): a brick is allocated iff
// |field(brick centre)| < 1.8 * brick extent, voxels store binary16 RNE of the
// field at voxel centres brick_min + (v * vs + vs/2), slots are numbered in
// brick order. Here the brick order is (L2 cell ascending, then z-y-x inside
// the cell) and the field is
//     field(p) = min over groups g of chain_g(p),
//     chain_g  = d_0; d = smin(d, d_i, k) over the group's primitives in order
// (sdf.rs:48-51 smin; evaluate_primitive voxelize_object.rs:439-474 after the
// primitive's translation and rotation).
//
// Exactness of the culling: primitive j is binned into every L2 cell its
// bounding sphere grown by M = 3 * brick + 2k touches, and a cell evaluates
// only the groups binned there. An unbinned group has every primitive at
// distance > M from the cell, so its chain exceeds M - 7k/4 > 3 bricks
// (smin lowers its inputs' minimum by at most k/4 per step), while the field
// at any voxel of an allocated brick is below 1.8 + 0.87 bricks: the culled
// evaluation equals the full one wherever it is stored or tested.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "ref_eval.h"

namespace rkf {

// Grids of one block per item: 2-D so that blocks x 512 lanes stay below the
// 2^32 work-items-per-dimension limit.
constexpr uint32_t kGridX = 32768u;
inline dim3 item_grid(uint64_t n) {
    uint32_t x = n < kGridX ? (uint32_t)n : kGridX;
    uint32_t y = (uint32_t)((n + x - 1) / x);
    return dim3(x, y, 1);
}
RKF_DEV uint64_t item_index() { return (uint64_t)blockIdx.y * gridDim.x + blockIdx.x; }

struct BuildDev {
    const rkfm_prim *prims;
    const uint32_t *group_prims;   // prim indices grouped, in original order
    const uint32_t *group_start;   // per group
    const uint32_t *group_count;
    const uint32_t *cand_cell;     // per candidate cell: flat L2 index
    const uint32_t *cand_off;      // CSR into cand_groups (n_cand + 1)
    const uint32_t *cand_groups;
    float origin[3];
    float vs;
    float k;
    uint32_t cells[3];
    uint64_t mat_seed;
};

RKF_DEV uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

RKF_DEV float box_sdf(float hx, float hy, float hz, float px, float py, float pz) {
    float qx = fabsf(px) - hx, qy = fabsf(py) - hy, qz = fabsf(pz) - hz;
    float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f), oz = fmaxf(qz, 0.0f);
    float outside = __builtin_sqrtf(ox * ox + oy * oy + oz * oz);
    float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
    return outside + inside;
}

// sdf.rs:21-37 capsule with a = (0,-hh,0), b = (0,hh,0).
RKF_DEV float capsule_sdf(float r, float hh, float px, float py, float pz) {
    float pax = px - 0.0f, pay = py - -hh, paz = pz - 0.0f;
    float bax = 0.0f - 0.0f, bay = hh - -hh, baz = 0.0f - 0.0f;
    float h = (pax * bax + pay * bay + paz * baz) / (bax * bax + bay * bay + baz * baz);
    h = clampf(h, 0.0f, 1.0f);
    float dx = pax - bax * h, dy = pay - bay * h, dz = paz - baz * h;
    return __builtin_sqrtf(dx * dx + dy * dy + dz * dz) - r;
}

// fBm terrain (prim type 3): value noise on a hashed integer lattice,
// smoothstep-interpolated, 6 octaves. The oracle evaluates the same
// operations in the same order.
RKF_DEV float terrain_lattice(uint32_t seed, int32_t i, int32_t j) {
    uint64_t key = (uint64_t)(uint32_t)i | ((uint64_t)(uint32_t)j << 32);
    uint64_t h = mix64(key ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull));
    return (float)(uint32_t)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}
RKF_DEV float terrain_vnoise(uint32_t seed, float u, float v) {
    float fu = floorf(u), fv = floorf(v);
    int32_t iu = (int32_t)fu, iv = (int32_t)fv;
    float tu = u - fu, tv = v - fv;
    float su = tu * tu * (3.0f - 2.0f * tu), sv = tv * tv * (3.0f - 2.0f * tv);
    float c00 = terrain_lattice(seed, iu, iv), c10 = terrain_lattice(seed, iu + 1, iv);
    float c01 = terrain_lattice(seed, iu, iv + 1), c11 = terrain_lattice(seed, iu + 1, iv + 1);
    float a = c00 + (c10 - c00) * su;
    float b = c01 + (c11 - c01) * su;
    return a + (b - a) * sv;
}
RKF_DEV float terrain_height(const rkfm_prim &pr, float x, float z) {
    float u = x - pr.centre[0], w = z - pr.centre[2];
    float h = 0.0f, amp = 1.0f, freq = pr.params[1];
    for (uint32_t i = 0; i < 6u; i++) {
        h = h + amp * terrain_vnoise(pr.seed + i, u * freq, w * freq);
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return h;
}
RKF_DEV float terrain_from_height(const rkfm_prim &pr, float y, float h) {
    return pr.params[2] * ((y - pr.centre[1]) - pr.params[0] * h);
}
RKF_DEV float terrain_sdf(const rkfm_prim &pr, float x, float y, float z) {
    return terrain_from_height(pr, y, terrain_height(pr, x, z));
}

RKF_DEV float prim_eval(const rkfm_prim &pr, float x, float y, float z) {
    if (pr.type == 3u) return terrain_sdf(pr, x, y, z);
    float d0 = x - pr.centre[0], d1 = y - pr.centre[1], d2 = z - pr.centre[2];
    const float *R = pr.rot;
    float lx = R[0] * d0 + R[1] * d1 + R[2] * d2;
    float ly = R[3] * d0 + R[4] * d1 + R[5] * d2;
    float lz = R[6] * d0 + R[7] * d1 + R[8] * d2;
    switch (pr.type) {
    case 0: return __builtin_sqrtf(lx * lx + ly * ly + lz * lz) - pr.params[0];
    case 1: return box_sdf(pr.params[0], pr.params[1], pr.params[2], lx, ly, lz);
    case 2: return capsule_sdf(pr.params[0], pr.params[1], lx, ly, lz);
    default: return kMaxFloat;
    }
}

// sdf.rs:48-51
RKF_DEV float smin(float a, float b, float k) {
    float h = fmaxf(k - fabsf(a - b), 0.0f) / k;
    return fminf(a, b) - h * h * k * 0.25f;
}

RKF_DEV float scene_field(const BuildDev &B, uint32_t cand, float x, float y, float z) {
    float d = INFINITY;
    uint32_t g0 = B.cand_off[cand], g1 = B.cand_off[cand + 1];
    for (uint32_t gi = g0; gi < g1; gi++) {
        uint32_t g = B.cand_groups[gi];
        uint32_t s = B.group_start[g], n = B.group_count[g];
        float r = prim_eval(B.prims[B.group_prims[s]], x, y, z);
        for (uint32_t j = 1; j < n; j++) {
            float dj = prim_eval(B.prims[B.group_prims[s + j]], x, y, z);
            r = B.k > 0.0f ? smin(r, dj, B.k) : fminf(r, dj);
        }
        d = fminf(d, r);
    }
    return d;
}

// scene_field with one terrain primitive's height given (prim index tp,
// height th at this point's column): the same operations as prim_eval.
RKF_DEV float scene_field_th(const BuildDev &B, uint32_t cand, float x, float y, float z, uint32_t tp, float th) {
    float d = INFINITY;
    uint32_t g0 = B.cand_off[cand], g1 = B.cand_off[cand + 1];
    for (uint32_t gi = g0; gi < g1; gi++) {
        uint32_t g = B.cand_groups[gi];
        uint32_t s = B.group_start[g], n = B.group_count[g];
        uint32_t pi = B.group_prims[s];
        float r = pi == tp ? terrain_from_height(B.prims[pi], y, th) : prim_eval(B.prims[pi], x, y, z);
        for (uint32_t j = 1; j < n; j++) {
            pi = B.group_prims[s + j];
            float dj = pi == tp ? terrain_from_height(B.prims[pi], y, th) : prim_eval(B.prims[pi], x, y, z);
            r = B.k > 0.0f ? smin(r, dj, B.k) : fminf(r, dj);
        }
        d = fminf(d, r);
    }
    return d;
}

// One block per candidate cell, one lane per brick (x + 8y + 64z in the cell).
__global__ __launch_bounds__(512) void alloc_flags_kernel(BuildDev B, uint32_t ncand, uint8_t *flags,
                                                          uint32_t *counts) {
    __shared__ uint32_t wsum[8];
    uint64_t item = item_index();
    if (item >= ncand) return;
    uint32_t cand = (uint32_t)item, t = threadIdx.x;
    uint32_t cell = B.cand_cell[cand];
    uint32_t qx = cell % B.cells[0], qy = (cell / B.cells[0]) % B.cells[1], qz = cell / (B.cells[0] * B.cells[1]);
    uint32_t gbx = qx * 8u + (t & 7u), gby = qy * 8u + ((t >> 3) & 7u), gbz = qz * 8u + (t >> 6);
    float bws = B.vs * 8.0f;
    float hb = bws * 0.5f;
    float cx = (B.origin[0] + (float)gbx * bws) + hb;
    float cy = (B.origin[1] + (float)gby * bws) + hb;
    float cz = (B.origin[2] + (float)gbz * bws) + hb;
    float d = scene_field(B, cand, cx, cy, cz);
    bool need = fabsf(d) < bws * 1.8f;
    flags[(uint64_t)cand * 512u + t] = need ? 1 : 0;
    uint64_t ball = __ballot(need);
    if ((t & 63u) == 0) wsum[t >> 6] = (uint32_t)__popcll(ball);
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int w = 0; w < 8; w++) s += wsum[w];
        counts[cand] = s;
    }
}

// Assign slots (base + rank in z-y-x order) and write L1/L2 entries.
__global__ __launch_bounds__(512) void assign_kernel(BuildDev B, uint32_t ncand, const uint8_t *flags,
                                                     const uint32_t *blk_of, const uint32_t *slot_base, uint32_t *l2,
                                                     uint32_t *l1, uint32_t *slot_cand, uint32_t *slot_brick) {
    __shared__ uint32_t wsum[8];
    uint64_t item = item_index();
    if (item >= ncand) return;
    uint32_t cand = (uint32_t)item, t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t blk = blk_of[cand];
    if (blk == RKF_EMPTY_SLOT) return;  // uniform per block
    bool need = flags[(uint64_t)cand * 512u + t] != 0;
    uint64_t ball = __ballot(need);
    if (lane == 0) wsum[w] = (uint32_t)__popcll(ball);
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < w; i++) before += wsum[i];
    uint32_t rank = before + (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
    uint32_t slot = slot_base[cand] + rank;
    l1[(uint64_t)blk * 512u + t] = need ? slot : RKF_EMPTY_SLOT;
    if (need) {
        slot_cand[slot] = cand;
        slot_brick[slot] = t;
    }
    if (t == 0) l2[B.cand_cell[cand]] = blk;
}

// One block per allocated brick, one lane per voxel (populate_brick). The
// first terrain primitive of the cell's groups has its height computed once
// per voxel column (64 per brick) and shared through LDS: the height does not
// depend on y.
__global__ __launch_bounds__(512) void fill_kernel(BuildDev B, uint64_t nbricks, const uint32_t *slot_cand,
                                                   const uint32_t *slot_brick, uint16_t *dist, uint8_t *mat) {
    __shared__ float hcol[64];
    __shared__ uint32_t tprim;
    uint64_t item = item_index();
    if (item >= nbricks) return;
    uint32_t slot = (uint32_t)item, v = threadIdx.x;
    uint32_t cand = slot_cand[slot], lb = slot_brick[slot];
    uint32_t cell = B.cand_cell[cand];
    uint32_t qx = cell % B.cells[0], qy = (cell / B.cells[0]) % B.cells[1], qz = cell / (B.cells[0] * B.cells[1]);
    uint32_t gbx = qx * 8u + (lb & 7u), gby = qy * 8u + ((lb >> 3) & 7u), gbz = qz * 8u + (lb >> 6);
    uint32_t vx = v & 7u, vy = (v >> 3) & 7u, vz = v >> 6;
    float vs = B.vs, bws = vs * 8.0f, half = vs * 0.5f;
    float px = (B.origin[0] + (float)gbx * bws) + ((float)vx * vs + half);
    float py = (B.origin[1] + (float)gby * bws) + ((float)vy * vs + half);
    float pz = (B.origin[2] + (float)gbz * bws) + ((float)vz * vs + half);
    if (v == 0u) {
        uint32_t tp = ~0u;
        for (uint32_t gi = B.cand_off[cand]; gi < B.cand_off[cand + 1] && tp == ~0u; gi++) {
            const uint32_t g = B.cand_groups[gi];
            for (uint32_t j = 0; j < B.group_count[g]; j++) {
                const uint32_t pi = B.group_prims[B.group_start[g] + j];
                if (B.prims[pi].type == 3u) {
                    tp = pi;
                    break;
                }
            }
        }
        tprim = tp;
    }
    __syncthreads();
    const uint32_t tp = tprim;
    if (tp != ~0u && v < 64u) {
        // column (v & 7, v >> 3): the x and z of its voxels, as above
        const float cxw = (B.origin[0] + (float)gbx * bws) + ((float)(v & 7u) * vs + half);
        const float czw = (B.origin[2] + (float)gbz * bws) + ((float)(v >> 3) * vs + half);
        hcol[v] = terrain_height(B.prims[tp], cxw, czw);
    }
    __syncthreads();
    float d = tp != ~0u ? scene_field_th(B, cand, px, py, pz, tp, hcol[vx + 8u * vz]) : scene_field(B, cand, px, py, pz);
    uint64_t idx = (uint64_t)slot * 512u + v;
    dist[idx] = (uint16_t)f2h(d);
    uint32_t m = 0;
    if (B.mat_seed) {
        uint64_t key = (uint64_t)(gbx * 8u + vx) | ((uint64_t)(gby * 8u + vy) << 21) | ((uint64_t)(gbz * 8u + vz) << 42);
        m = (uint32_t)(mix64(B.mat_seed ^ mix64(key ^ 0ull)) & 15u);
    }
    mat[idx] = (uint8_t)m;
}

// Free masks: one block per row cell, one lane per brick b of the cell;
// bit set = some brick of {b..b+1}^3 (clamped to the grid) is allocated.
__global__ __launch_bounds__(512) void free_mask_kernel(SparseDev S, uint64_t nrows, const uint32_t *row_cell,
                                                        uint64_t *bits, uint32_t *row_any) {
    __shared__ uint32_t any;
    uint64_t item = item_index();
    if (item >= nrows) return;
    uint32_t row = (uint32_t)item, t = threadIdx.x;
    if (t == 0) any = 0;
    __syncthreads();
    uint32_t cell = row_cell[row];
    uint32_t qx = cell % S.cells[0], qy = (cell / S.cells[0]) % S.cells[1], qz = cell / (S.cells[0] * S.cells[1]);
    uint32_t bx = qx * 8u + (t & 7u), by = qy * 8u + ((t >> 3) & 7u), bz = qz * 8u + (t >> 6);
    bool occupied = false;
    for (uint32_t dz = 0; dz < 2; dz++)
        for (uint32_t dy = 0; dy < 2; dy++)
            for (uint32_t dx = 0; dx < 2; dx++) {
                uint32_t x = min(bx + dx, S.nb[0] - 1), y = min(by + dy, S.nb[1] - 1), z = min(bz + dz, S.nb[2] - 1);
                uint64_t c = (uint64_t)(x >> 3) + (uint64_t)(y >> 3) * S.cells[0] +
                             (uint64_t)(z >> 3) * S.cells[0] * S.cells[1];
                uint32_t blk = S.l2[c];
                if (blk == RKF_EMPTY_SLOT) continue;
                uint32_t s = S.l1[(uint64_t)blk * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u];
                if (s != RKF_EMPTY_SLOT) occupied = true;
            }
    uint64_t ball = __ballot(occupied);
    if ((t & 63u) == 0) {
        bits[(uint64_t)row * 8u + (t >> 6)] = ball;
        if (ball) atomicOr(&any, 1u);
    }
    __syncthreads();
    if (t == 0) row_any[row] = any;
}

// Octant extents. For octant o (bit a set = axis a runs negative) and a cell
// C: E_o(C) = 0 if C is not entirely free, else min(15, 1 + min over the 7
// neighbours C + (i sx, j sy, k sz), (i, j, k) in {0,1}^3 minus 0), with
// out-of-grid neighbours counted as 15 (free): the side of the largest cube of
// free cells that has C at its corner and extends in o's directions. Packed
// 4 bits per octant. Jacobi passes from "free = 15" reach the fixed point in
// 15 passes (a value below 15 is the octant distance to a non-free cell).
__global__ __launch_bounds__(256) void oct_init_kernel(const uint32_t *free_idx, uint64_t n, uint32_t *oct) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) oct[i] = free_idx[i] == kFreeCell ? 0xFFFFFFFFu : 0u;
}
__global__ __launch_bounds__(256) void oct_relax_kernel(uint32_t cx, uint32_t cy, uint32_t cz, const uint32_t *cur,
                                                        uint32_t *nxt) {
    uint64_t n = (uint64_t)cx * cy * cz;
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    if (cur[i] == 0u) {
        nxt[i] = 0u;
        return;
    }
    uint32_t x = (uint32_t)(i % cx), y = (uint32_t)((i / cx) % cy), z = (uint32_t)(i / ((uint64_t)cx * cy));
    uint32_t nb[27];
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int64_t xx = (int64_t)x + dx, yy = (int64_t)y + dy, zz = (int64_t)z + dz;
                bool in = xx >= 0 && yy >= 0 && zz >= 0 && xx < cx && yy < cy && zz < cz;
                nb[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] =
                    in ? cur[(uint64_t)xx + (uint64_t)yy * cx + (uint64_t)zz * cx * cy] : 0xFFFFFFFFu;
            }
    uint32_t out = 0;
    for (int o = 0; o < 8; o++) {
        int sx = (o & 1) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 4) ? -1 : 1;
        uint32_t m = kOctCap;
        for (int k = 1; k < 8; k++) {
            int ox = (k & 1) * sx, oy = ((k >> 1) & 1) * sy, oz = (k >> 2) * sz;
            uint32_t v = (nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)] >> (4 * o)) & 15u;
            m = v < m ? v : m;
        }
        uint32_t e = m + 1u < kOctCap ? m + 1u : kOctCap;
        out |= e << (4 * o);
    }
    nxt[i] = out;
}
constexpr uint32_t kOctPasses = 16u;  // >= 15, even: the result lands in the first buffer

// Brick octant extents inside rows (cells that are not entirely free): for
// a brick b of a row and octant o, E_o(b) = 0 if b is not free, else
// min(15, 1 + min over the 7 neighbours b + (i sx, j sy, k sz)) where a
// neighbour outside the brick grid counts 15 and a neighbour n in an entirely
// free cell C counts the bound its cell extent gives, min(15, distance from n
// to the far side of the free cube of cells C + [0, E_o(C)) on each axis),
// which is at least 1. Every brick of the cube b + [0, E_o(b)) in o's
// directions is then free. Jacobi passes from "free = 15" over packed u32
// (8 nibbles per brick) reach the fixed point in 15 passes; the march reads
// them from the row words.
struct BextArgs {
    uint32_t cells[3];
    uint32_t nb[3];
    const uint32_t *free_idx;
    const uint32_t *oct;
    const uint64_t *free_bits;
    const uint32_t *row_cell;
    uint64_t n_rows;
};

RKF_DEV uint32_t free_cell_bound(const BextArgs &A, uint32_t cell, int64_t x, int64_t y, int64_t z) {
    const uint32_t ov = A.oct[cell];
    const int64_t c[3] = {(int64_t)(cell % A.cells[0]), (int64_t)((cell / A.cells[0]) % A.cells[1]),
                          (int64_t)(cell / (A.cells[0] * A.cells[1]))};
    const int64_t n[3] = {x, y, z};
    uint32_t out = 0;
    for (int o = 0; o < 8; o++) {
        const int64_t e = (ov >> (4 * o)) & 15u;
        int64_t m = kOctCap;
        for (int a = 0; a < 3; a++) {
            int64_t d = ((o >> a) & 1) ? n[a] - (c[a] - e + 1) * 8 + 1 : (c[a] + e) * 8 - n[a];
            m = d < m ? d : m;
        }
        out |= (uint32_t)m << (4 * o);
    }
    return out;
}

__global__ __launch_bounds__(512) void bext_init_kernel(BextArgs A, uint32_t *cur) {
    uint64_t item = item_index();
    if (item >= A.n_rows) return;
    const uint32_t t = threadIdx.x;
    bool nf = (A.free_bits[item * 8u + (t >> 6)] >> (t & 63u)) & 1ull;
    cur[item * 512u + t] = nf ? 0u : 0xFFFFFFFFu;
}

// One Jacobi pass, one workgroup per row: the row's 8x8x8 bricks and the
// one-brick shell around them (from neighbouring rows, the free-cell bound
// of free neighbouring cells, or 15 past the grid end) are staged in LDS, so
// each brick's 26 neighbours are LDS reads.
constexpr uint32_t kOutOfGrid = 0xFFFFFFFEu;

__global__ __launch_bounds__(512) void bext_relax_kernel(BextArgs A, const uint32_t *cur, uint32_t *nxt) {
    __shared__ uint32_t tile[1000];  // 10 x 10 x 10 bricks: the cell and its shell
    __shared__ uint32_t nrow[27];    // rows of the 3 x 3 x 3 cells around (kFreeCell / kOutOfGrid)
    const uint64_t item = item_index();
    if (item >= A.n_rows) return;
    const uint32_t t = threadIdx.x;
    const uint32_t cell = A.row_cell[item];
    const int64_t cq[3] = {(int64_t)(cell % A.cells[0]), (int64_t)((cell / A.cells[0]) % A.cells[1]),
                           (int64_t)(cell / (A.cells[0] * A.cells[1]))};
    if (t < 27u) {
        const int64_t x = cq[0] + (int64_t)(t % 3u) - 1, y = cq[1] + (int64_t)((t / 3u) % 3u) - 1,
                      z = cq[2] + (int64_t)(t / 9u) - 1;
        nrow[t] = (x < 0 || y < 0 || z < 0 || x >= A.cells[0] || y >= A.cells[1] || z >= A.cells[2])
                      ? kOutOfGrid
                      : (t == 13u ? (uint32_t)item
                                  : A.free_idx[(uint64_t)x + (uint64_t)y * A.cells[0] + (uint64_t)z * A.cells[0] * A.cells[1]]);
    }
    __syncthreads();
    for (uint32_t i = t; i < 1000u; i += 512u) {
        const uint32_t h[3] = {i % 10u, (i / 10u) % 10u, i / 100u};
        int64_t bb[3];
        uint32_t nc = 0u;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            bb[a] = cq[a] * 8 - 1 + (int64_t)h[a];
            nc += (h[a] == 0u ? 0u : (h[a] == 9u ? 2u : 1u)) * (a == 0 ? 1u : (a == 1 ? 3u : 9u));
        }
        const uint32_t r2 = nrow[nc];
        uint32_t v;
        if (r2 == kOutOfGrid) {
            v = 0xFFFFFFFFu;
        } else if (r2 == kFreeCell) {
            const uint32_t c2 = (uint32_t)((bb[0] >> 3) + (bb[1] >> 3) * A.cells[0] + (bb[2] >> 3) * A.cells[0] * A.cells[1]);
            v = free_cell_bound(A, c2, bb[0], bb[1], bb[2]);
        } else {
            v = cur[(uint64_t)r2 * 512u + (uint32_t)((bb[0] & 7) + (bb[1] & 7) * 8 + (bb[2] & 7) * 64)];
        }
        tile[i] = v;
    }
    __syncthreads();
    const uint32_t x = t & 7u, y = (t >> 3) & 7u, z = t >> 6;
    const uint32_t c = (z + 1u) * 100u + (y + 1u) * 10u + (x + 1u);
    if (tile[c] == 0u) {
        nxt[item * 512u + t] = 0u;
        return;
    }
    uint32_t out = 0;
    for (int o = 0; o < 8; o++) {
        const int sx = (o & 1) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 4) ? -1 : 1;
        uint32_t m = kOctCap;
        for (int k = 1; k < 8; k++) {
            const int ox = (k & 1) * sx, oy = ((k >> 1) & 1) * sy, oz = (k >> 2) * sz;
            const uint32_t v = (tile[(int)c + ox + 10 * oy + 100 * oz] >> (4 * o)) & 15u;
            m = v < m ? v : m;
        }
        const uint32_t e = m + 1u < kOctCap ? m + 1u : kOctCap;
        out |= e << (4 * o);
    }
    nxt[item * 512u + t] = out;
}

// Row word per brick: packed octant extents << 32 | L1 slot (EMPTY in rows
// past n_blocks, which hold no block): one load gives a DDA step both the
// brick's extent and the slot it samples.
__global__ __launch_bounds__(256) void rowword_kernel(uint64_t n_rows, uint32_t n_blocks, const uint32_t *packed,
                                                      const uint32_t *l1, uint64_t *roww) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x + (uint64_t)blockIdx.y * gridDim.x * 256u;
    if (i >= n_rows * 512u) return;
    const uint64_t row = i / 512u;
    const uint32_t slot = row < n_blocks ? l1[i] : RKF_EMPTY_SLOT;
    roww[i] = ((uint64_t)packed[i] << 32) | slot;
}

// cellw = free row << 32 | octant extents (row kFreeCell: every brick free).
__global__ __launch_bounds__(256) void cellword_kernel(const uint32_t *free_idx, const uint32_t *oct, uint64_t n,
                                                       uint64_t *cellw) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    cellw[i] = ((uint64_t)free_idx[i] << 32) | (uint64_t)oct[i];
}

// ---- reference scene -> sparse scene (rkfm_sparse_from_objects) ----
// The field is the minimum of evaluate_object (ray_march.wgsl:600-634) over
// the candidate objects of the point's L2 cell (the objects whose AABB,
// widened by 3 bricks, meets the cell), in ascending object index; the first
// object reaching the minimum gives the material (evaluate_object's) and the
// aux word (secondary material | blend weight << 8 of its nearest voxel,
// :570-596).
struct ObjBuildDev {
    RefParams R;                // pool, maps, objects, su.num_objects
    const uint32_t *cand_cell;  // per candidate cell: flat L2 index
    const uint32_t *cand_off;   // CSR into cand_objs (n_cand + 1)
    const uint32_t *cand_objs;
    float origin[3];
    float vs;
    uint32_t cells[3];
};

RKF_DEV float objects_field(const ObjBuildDev &B, uint32_t cand, f3 p, uint32_t &win, uint32_t &mat) {
    RefCtx c{&B.R, 0u};
    float best = INFINITY;
    win = ~0u;
    mat = 0u;
    for (uint32_t k = B.cand_off[cand]; k < B.cand_off[cand + 1]; k++) {
        const uint32_t oi = B.cand_objs[k];
        uint32_t m;
        const float d = eval_object(c, p, oi, m);
        if (d < best) {
            best = d;
            win = oi;
            mat = m;
        }
    }
    return best;
}

// One block per candidate cell, one lane per brick: stored iff |field at the
// brick centre| < 1.8 brick extents (voxelize_object.rs:86,104).
__global__ __launch_bounds__(512) void obj_alloc_kernel(ObjBuildDev B, uint32_t ncand, uint8_t *flags,
                                                        uint32_t *counts) {
    __shared__ uint32_t wsum[8];
    uint64_t item = item_index();
    if (item >= ncand) return;
    uint32_t cand = (uint32_t)item, t = threadIdx.x;
    uint32_t cell = B.cand_cell[cand];
    uint32_t qx = cell % B.cells[0], qy = (cell / B.cells[0]) % B.cells[1], qz = cell / (B.cells[0] * B.cells[1]);
    uint32_t gbx = qx * 8u + (t & 7u), gby = qy * 8u + ((t >> 3) & 7u), gbz = qz * 8u + (t >> 6);
    float bws = B.vs * 8.0f;
    float hb = bws * 0.5f;
    f3 c = mk((B.origin[0] + (float)gbx * bws) + hb, (B.origin[1] + (float)gby * bws) + hb,
              (B.origin[2] + (float)gbz * bws) + hb);
    uint32_t win, m;
    float d = objects_field(B, cand, c, win, m);
    bool need = fabsf(d) < bws * 1.8f;
    flags[(uint64_t)cand * 512u + t] = need ? 1 : 0;
    uint64_t ball = __ballot(need);
    if ((t & 63u) == 0) wsum[t >> 6] = (uint32_t)__popcll(ball);
    __syncthreads();
    if (t == 0) {
        uint32_t n = 0;
        for (int w = 0; w < 8; w++) n += wsum[w];
        counts[cand] = n;
    }
}

// One block per stored brick, one lane per voxel (centre brick_min + (v + 0.5) vs).
__global__ __launch_bounds__(512) void obj_fill_kernel(ObjBuildDev B, uint64_t nbricks, const uint32_t *slot_cand,
                                                       const uint32_t *slot_brick, uint16_t *dist, uint8_t *mat,
                                                       uint16_t *aux, uint32_t *any_aux) {
    uint64_t item = item_index();
    if (item >= nbricks) return;
    uint32_t slot = (uint32_t)item, v = threadIdx.x;
    uint32_t cand = slot_cand[slot], lb = slot_brick[slot];
    uint32_t cell = B.cand_cell[cand];
    uint32_t qx = cell % B.cells[0], qy = (cell / B.cells[0]) % B.cells[1], qz = cell / (B.cells[0] * B.cells[1]);
    uint32_t gbx = qx * 8u + (lb & 7u), gby = qy * 8u + ((lb >> 3) & 7u), gbz = qz * 8u + (lb >> 6);
    uint32_t vx = v & 7u, vy = (v >> 3) & 7u, vz = v >> 6;
    float vs = B.vs, bws = vs * 8.0f, half = vs * 0.5f;
    f3 p = mk((B.origin[0] + (float)gbx * bws) + ((float)vx * vs + half),
              (B.origin[1] + (float)gby * bws) + ((float)vy * vs + half),
              (B.origin[2] + (float)gbz * bws) + ((float)vz * vs + half));
    uint32_t win, m;
    float d = objects_field(B, cand, p, win, m);
    uint64_t idx = (uint64_t)slot * 512u + v;
    dist[idx] = (uint16_t)f2h(d);
    mat[idx] = (uint8_t)(m & 0x3Fu);
    uint32_t a = 0;
    if (win != ~0u && B.R.objects[win].sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        const rkf_gpu_object &o = B.R.objects[win];
        uint32_t sl, vox;
        uint64_t i = nearest_voxel(B.R, xform_point(o.inverse_world, p), o, sl, vox);
        if (i != ~0ull) a = ((pool_w0(B.R, i) >> 22) & 0x3Fu) | (((pool_w1(B.R, i) >> 24) & 0xFFu) << 8);
    }
    aux[idx] = (uint16_t)a;
    if (a) atomicOr(any_aux, 1u);
}

}  // namespace rkf

rkf::SparseDev rkfm_sparse_scene::dev() const {
    rkf::SparseDev d;
    memset(&d, 0, sizeof(d));
    for (int a = 0; a < 3; a++) {
        d.origin[a] = info.origin[a];
        d.cells[a] = info.cells[a];
        d.nb[a] = info.cells[a] * 8u;
        d.tot[a] = (int32_t)(info.cells[a] * 64u);
    }
    d.vs = info.voxel_size;
    d.inv_vs = 1.0f / info.voxel_size;
    d.l2 = (const uint32_t *)l2.ptr;
    d.l1 = (const uint32_t *)l1.ptr;
    d.dist = (const uint16_t *)dist.ptr;
    d.mat = (const uint8_t *)mat.ptr;
    d.aux = (const uint16_t *)aux.ptr;
    d.free_idx = (const uint32_t *)free_idx.ptr;
    d.free_bits = (const uint64_t *)free_bits.ptr;
    d.oct = (const uint32_t *)oct.ptr;
    d.roww = (const uint64_t *)roww.ptr;
    d.cellw = (const uint64_t *)cellw.ptr;
    d.n_cells = (uint64_t)info.cells[0] * info.cells[1] * info.cells[2];
    d.n_blocks = info.n_blocks;
    d.block_cell = (const uint32_t *)block_cell.ptr;
    return d;
}

static void release_scene(rkfm_sparse_scene *s) {
    s->l2.release();
    s->l1.release();
    s->dist.release();
    s->mat.release();
    s->aux.release();
    s->free_idx.release();
    s->free_bits.release();
    s->oct.release();
    s->cellw.release();
    s->roww.release();
    s->block_cell.release();
    s->edit_touched.release();
}

void rkf_sparse_release_all(rkfm_ctx *c) {
    for (rkfm_sparse_scene *s : c->scenes) {
        release_scene(s);
        delete s;
    }
    c->scenes.clear();
}

namespace {

struct TmpBufs {
    std::vector<DevBuf *> bufs;
    ~TmpBufs() {
        for (DevBuf *b : bufs) {
            b->release();
            delete b;
        }
    }
    DevBuf *make() {
        bufs.push_back(new DevBuf());
        return bufs.back();
    }
};

int put(hipStream_t st, DevBuf &b, const void *src, uint64_t bytes) {
    int rc = b.reserve(bytes + 1024, false);  // kernels read the first KiB and 4-byte words at the end
    if (rc) return rc;
    b.size = bytes;
    if (bytes) rc = rkf_hip_check(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st), "upload");
    return rc;
}

uint64_t used_bytes(const rkfm_ctx *c) {
    uint64_t s = 0;
    for (const rkfm_sparse_scene *x : c->scenes) s += x->info.device_bytes;
    return s;
}

// Builds free_idx / free_bits from l2/l1 already on the device.
int build_free_masks(rkfm_ctx *c, rkfm_sparse_scene *s, const std::vector<uint32_t> &block_cells) {
    const uint32_t *cells = s->info.cells;
    uint64_t ncell = (uint64_t)cells[0] * cells[1] * cells[2];
    // Rows: cells whose {C..C+1}^3 neighbourhood holds a block, i.e. C in
    // block - {0,1}^3. Row i < n_blocks is block i's own cell (an occupied
    // cell is never entirely free), so a row number is also the cell's L1
    // block when it is below n_blocks.
    std::vector<uint32_t> rows(block_cells), extra;
    std::vector<uint8_t> occupied(ncell, 0);
    for (uint32_t bc : block_cells) occupied[bc] = 1;
    extra.reserve(block_cells.size() * 7);
    for (uint32_t bc : block_cells) {
        uint32_t qx = bc % cells[0], qy = (bc / cells[0]) % cells[1], qz = bc / (cells[0] * cells[1]);
        for (uint32_t dz = 0; dz < 2; dz++)
            for (uint32_t dy = 0; dy < 2; dy++)
                for (uint32_t dx = 0; dx < 2; dx++) {
                    if (qx < dx || qy < dy || qz < dz) continue;
                    uint32_t cc = (qx - dx) + (qy - dy) * cells[0] + (qz - dz) * cells[0] * cells[1];
                    if (!occupied[cc]) extra.push_back(cc);
                }
    }
    std::sort(extra.begin(), extra.end());
    extra.erase(std::unique(extra.begin(), extra.end()), extra.end());
    rows.insert(rows.end(), extra.begin(), extra.end());
    if (rows.size() >= (uint64_t)rkf::kFreeCell)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "sparse scene: %zu free-mask rows", rows.size());
    TmpBufs tmp;
    DevBuf *d_rows = tmp.make(), *d_any = tmp.make();
    int rc = put(c->stream, s->block_cell, block_cells.data(), block_cells.size() * 4);
    if (!rc) rc = put(c->stream, *d_rows, rows.data(), rows.size() * 4);
    if (!rc) rc = d_any->reserve(rows.size() * 4 + 4, false);
    if (!rc) rc = s->free_bits.reserve(rows.size() * 64 + 64, false);
    if (!rc) rc = s->free_idx.reserve(ncell * 4, false);
    if (rc) return rc;
    s->free_bits.size = rows.size() * 64;
    s->free_idx.size = ncell * 4;
    std::vector<uint32_t> any(rows.size());
    if (!rows.empty()) {
        rkf::SparseDev D = s->dev();
        hipLaunchKernelGGL(rkf::free_mask_kernel, rkf::item_grid(rows.size()), dim3(512), 0, c->stream, D,
                           (uint64_t)rows.size(), (const uint32_t *)d_rows->ptr, (uint64_t *)s->free_bits.ptr, (uint32_t *)d_any->ptr);
        rc = rkf_hip_check(hipGetLastError(), "free_mask launch");
        if (!rc)
            rc = rkf_hip_check(hipMemcpyAsync(any.data(), d_any->ptr, rows.size() * 4, hipMemcpyDeviceToHost, c->stream),
                               "free rows download");
        if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "free rows sync");
        if (rc) return rc;
    }
    std::vector<uint32_t> fidx(ncell, rkf::kFreeCell);
    for (size_t r = 0; r < rows.size(); r++)
        if (any[r]) fidx[rows[r]] = (uint32_t)r;
    s->n_free_rows = rows.size();
    rc = rkf_hip_check(hipMemcpyAsync(s->free_idx.ptr, fidx.data(), ncell * 4, hipMemcpyHostToDevice, c->stream),
                       "free idx upload");
    if (rc) return rc;
    // octant extents
    DevBuf *tmpb = tmp.make();
    rc = s->oct.reserve(ncell * 4, false);
    if (!rc) rc = tmpb->reserve(ncell * 4, false);
    if (rc) return rc;
    s->oct.size = ncell * 4;
    dim3 g((unsigned)((ncell + 255) / 256)), b(256);
    hipLaunchKernelGGL(rkf::oct_init_kernel, g, b, 0, c->stream, (const uint32_t *)s->free_idx.ptr, ncell,
                       (uint32_t *)s->oct.ptr);
    uint32_t *cur = (uint32_t *)s->oct.ptr, *nxt = (uint32_t *)tmpb->ptr;
    for (uint32_t pass = 0; pass < rkf::kOctPasses; pass++) {
        hipLaunchKernelGGL(rkf::oct_relax_kernel, g, b, 0, c->stream, cells[0], cells[1], cells[2], cur, nxt);
        std::swap(cur, nxt);
    }
    static_assert(rkf::kOctPasses % 2u == 0u && rkf::kOctPasses >= rkf::kOctCap, "pass count");
    rc = rkf_hip_check(hipGetLastError(), "octant launch");
    if (!rc) rc = s->cellw.reserve(ncell * 8, false);
    if (rc) return rc;
    s->cellw.size = ncell * 8;
    hipLaunchKernelGGL(rkf::cellword_kernel, g, b, 0, c->stream, (const uint32_t *)s->free_idx.ptr,
                       (const uint32_t *)s->oct.ptr, ncell, (uint64_t *)s->cellw.ptr);
    rc = rkf_hip_check(hipGetLastError(), "cell word launch");
    if (rc) return rc;
    // brick extents inside rows
    const uint64_t nrows = rows.size();
    if (!rc) rc = s->roww.reserve(nrows * 4096 + 64, false);
    DevBuf *pa = tmp.make(), *pb = tmp.make();
    if (!rc) rc = pa->reserve(nrows * 2048 + 64, false);
    if (!rc) rc = pb->reserve(nrows * 2048 + 64, false);
    if (rc) return rc;
    s->roww.size = nrows * 4096;
    if (nrows) {
        rkf::BextArgs A;
        for (int a = 0; a < 3; a++) {
            A.cells[a] = cells[a];
            A.nb[a] = cells[a] * 8u;
        }
        A.free_idx = (const uint32_t *)s->free_idx.ptr;
        A.oct = (const uint32_t *)s->oct.ptr;
        A.free_bits = (const uint64_t *)s->free_bits.ptr;
        A.row_cell = (const uint32_t *)d_rows->ptr;
        A.n_rows = nrows;
        hipLaunchKernelGGL(rkf::bext_init_kernel, rkf::item_grid(nrows), dim3(512), 0, c->stream, A,
                           (uint32_t *)pa->ptr);
        uint32_t *bc = (uint32_t *)pa->ptr, *bn = (uint32_t *)pb->ptr;
        for (uint32_t pass = 0; pass < rkf::kOctPasses; pass++) {
            hipLaunchKernelGGL(rkf::bext_relax_kernel, rkf::item_grid(nrows), dim3(512), 0, c->stream, A, bc, bn);
            std::swap(bc, bn);
        }
        const uint64_t n_w = nrows * 512u;
        const uint64_t blocks = (n_w + 255u) / 256u;
        dim3 pg((unsigned)(blocks < 65535u ? blocks : 65535u), (unsigned)((blocks + 65534u) / 65535u));
        hipLaunchKernelGGL(rkf::rowword_kernel, pg, dim3(256), 0, c->stream, nrows, s->info.n_blocks, bc,
                           (const uint32_t *)s->l1.ptr, (uint64_t *)s->roww.ptr);
        rc = rkf_hip_check(hipGetLastError(), "brick extent launch");
        if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "brick extent sync");
    }
    if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "free idx sync");
    return rc;
}

void finish_info(rkfm_sparse_scene *s) {
    s->info.device_bytes = s->l2.cap + s->l1.cap + s->dist.cap + s->mat.cap + s->aux.cap + s->free_idx.cap +
                           s->free_bits.cap + s->oct.cap + s->cellw.cap + s->roww.cap + s->block_cell.cap;
}

}  // namespace

extern "C" {

int rkfm_sparse_build(rkfm_ctx *c, const rkfm_prim *prims, uint32_t n, const rkfm_sparse_params *prm,
                      rkfm_sparse_scene **out) {
    if (!c || !prims || !prm || !out || n == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: argument");
    *out = nullptr;
    const float vs = prm->voxel_size, k = prm->smin_k;
    if (!(vs > 0.0f) || k < 0.0f) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: voxel size / k");
    const float bws = vs * 8.0f, cws = bws * 8.0f;
    uint32_t cells[3];
    for (int a = 0; a < 3; a++) {
        float ext = prm->bounds_max[a] - prm->bounds_min[a];
        if (!(ext > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: empty bounds");
        double cc = std::ceil((double)ext / (double)cws);
        if (cc > 32768.0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: grid too large");
        cells[a] = std::max(1u, (uint32_t)cc);
    }
    uint64_t ncell = (uint64_t)cells[0] * cells[1] * cells[2];
    if (ncell > 0xFFFFFFF0ull) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: too many cells");
    // Groups: ids compacted in ascending order, prims in original order.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) {
        if (prims[i].type > 3) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: prim %u type", i);
        if (prims[i].type == 3 && !(prims[i].params[2] > 0.0f))
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_build: terrain %u distance scale", i);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return prims[a].group < prims[b].group; });
    std::vector<uint32_t> gstart, gcount, gid_of(n);
    for (uint32_t i = 0; i < n; i++) {
        if (i == 0 || prims[order[i]].group != prims[order[i - 1]].group) {
            gstart.push_back(i);
            gcount.push_back(0);
        }
        gcount.back()++;
        gid_of[order[i]] = (uint32_t)gstart.size() - 1;
    }
    // Bin groups into cells.
    const float M = 3.0f * bws + 2.0f * k;
    std::vector<uint64_t> pairs;  // (cell << 32) | group
    for (uint32_t i = 0; i < n; i++) {
        const rkfm_prim &p = prims[i];
        if (p.type == 3) {
            // Terrain: |h| <= 1.96875 |p0|. Its group is binned into every
            // cell that is not entirely above y = centre.y + H + M / p2 (above
            // that the terrain's distance exceeds M); cells below keep it so
            // that deep ground stays negative under partly buried groups.
            double top = (double)p.centre[1] + 1.96875 * std::fabs((double)p.params[0]) + (double)M / p.params[2] + 1e-3;
            int64_t yhi = (int64_t)std::floor((top - prm->bounds_min[1]) / cws);
            yhi = std::min<int64_t>(yhi, (int64_t)cells[1] - 1);
            for (int64_t z = 0; z < (int64_t)cells[2]; z++)
                for (int64_t y = 0; y <= yhi; y++)
                    for (int64_t x = 0; x < (int64_t)cells[0]; x++) {
                        uint64_t cell = (uint64_t)x + (uint64_t)y * cells[0] + (uint64_t)z * cells[0] * cells[1];
                        pairs.push_back((cell << 32) | gid_of[i]);
                    }
            continue;
        }
        float R = p.type == 0 ? p.params[0]
                  : p.type == 1 ? std::sqrt(p.params[0] * p.params[0] + p.params[1] * p.params[1] + p.params[2] * p.params[2])
                                : p.params[0] + p.params[1];
        double E = (double)R + M + 1e-3;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)std::floor(((double)p.centre[a] - E - prm->bounds_min[a]) / cws);
            hi[a] = (int64_t)std::floor(((double)p.centre[a] + E - prm->bounds_min[a]) / cws);
            lo[a] = std::max<int64_t>(lo[a], 0);
            hi[a] = std::min<int64_t>(hi[a], (int64_t)cells[a] - 1);
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    double d2 = 0;
                    int64_t q[3] = {x, y, z};
                    for (int a = 0; a < 3; a++) {
                        double c0 = prm->bounds_min[a] + (double)q[a] * cws, c1 = c0 + cws;
                        double v = std::max(std::max(c0 - p.centre[a], (double)p.centre[a] - c1), 0.0);
                        d2 += v * v;
                    }
                    if (d2 > E * E) continue;
                    uint64_t cell = (uint64_t)x + (uint64_t)y * cells[0] + (uint64_t)z * cells[0] * cells[1];
                    pairs.push_back((cell << 32) | gid_of[i]);
                }
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::vector<uint32_t> cand_cell, cand_off, cand_groups;
    for (size_t i = 0; i < pairs.size(); i++) {
        uint32_t cell = (uint32_t)(pairs[i] >> 32);
        if (cand_cell.empty() || cand_cell.back() != cell) {
            cand_cell.push_back(cell);
            cand_off.push_back((uint32_t)cand_groups.size());
        }
        cand_groups.push_back((uint32_t)(pairs[i] & 0xFFFFFFFFu));
    }
    cand_off.push_back((uint32_t)cand_groups.size());
    const uint32_t ncand = (uint32_t)cand_cell.size();

    rkfm_sparse_scene *s = new rkfm_sparse_scene();
    s->owner = c;
    for (int a = 0; a < 3; a++) {
        s->info.origin[a] = prm->bounds_min[a];
        s->info.cells[a] = cells[a];
    }
    s->info.voxel_size = vs;
    TmpBufs tmp;
    DevBuf *d_prims = tmp.make(), *d_gp = tmp.make(), *d_gs = tmp.make(), *d_gc = tmp.make();
    DevBuf *d_cc = tmp.make(), *d_co = tmp.make(), *d_cg = tmp.make(), *d_flags = tmp.make(), *d_cnt = tmp.make();
    DevBuf *d_blk = tmp.make(), *d_base = tmp.make(), *d_scand = tmp.make(), *d_sbrick = tmp.make();
    auto fail = [&](int rc) {
        release_scene(s);
        delete s;
        return rc;
    };
    int rc = put(c->stream, *d_prims, prims, (uint64_t)n * sizeof(rkfm_prim));
    if (!rc) rc = put(c->stream, *d_gp, order.data(), (uint64_t)n * 4);
    if (!rc) rc = put(c->stream, *d_gs, gstart.data(), gstart.size() * 4);
    if (!rc) rc = put(c->stream, *d_gc, gcount.data(), gcount.size() * 4);
    if (!rc) rc = put(c->stream, *d_cc, cand_cell.data(), cand_cell.size() * 4);
    if (!rc) rc = put(c->stream, *d_co, cand_off.data(), cand_off.size() * 4);
    if (!rc) rc = put(c->stream, *d_cg, cand_groups.data(), cand_groups.size() * 4);
    if (!rc) rc = d_flags->reserve((uint64_t)ncand * 512 + 4, false);
    if (!rc) rc = d_cnt->reserve((uint64_t)ncand * 4 + 4, false);
    if (rc) return fail(rc);
    rkf::BuildDev B;
    memset(&B, 0, sizeof(B));
    B.prims = (const rkfm_prim *)d_prims->ptr;
    B.group_prims = (const uint32_t *)d_gp->ptr;
    B.group_start = (const uint32_t *)d_gs->ptr;
    B.group_count = (const uint32_t *)d_gc->ptr;
    B.cand_cell = (const uint32_t *)d_cc->ptr;
    B.cand_off = (const uint32_t *)d_co->ptr;
    B.cand_groups = (const uint32_t *)d_cg->ptr;
    for (int a = 0; a < 3; a++) {
        B.origin[a] = prm->bounds_min[a];
        B.cells[a] = cells[a];
    }
    B.vs = vs;
    B.k = k;
    B.mat_seed = prm->mat_seed;
    std::vector<uint32_t> counts(ncand);
    if (ncand) {
        hipLaunchKernelGGL(rkf::alloc_flags_kernel, rkf::item_grid(ncand), dim3(512), 0, c->stream, B, ncand,
                           (uint8_t *)d_flags->ptr,
                           (uint32_t *)d_cnt->ptr);
        rc = rkf_hip_check(hipGetLastError(), "alloc_flags launch");
        if (!rc)
            rc = rkf_hip_check(hipMemcpyAsync(counts.data(), d_cnt->ptr, (uint64_t)ncand * 4, hipMemcpyDeviceToHost,
                                              c->stream),
                               "counts download");
        if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "counts sync");
        if (rc) return fail(rc);
    }
    std::vector<uint32_t> blk_of(ncand, RKF_EMPTY_SLOT), base(ncand, 0), block_cells;
    uint64_t nbricks = 0;
    uint32_t nblocks = 0;
    for (uint32_t i = 0; i < ncand; i++) {
        if (!counts[i]) continue;
        blk_of[i] = nblocks++;
        block_cells.push_back(cand_cell[i]);
        base[i] = (uint32_t)nbricks;
        nbricks += counts[i];
    }
    if (nbricks >= 0xFFFFFFF0ull) return fail(rkf_set_error(RKFM_ERR_OOM, "sparse_build: %llu bricks exceed u32 slots",
                                                            (unsigned long long)nbricks));
    if (prm->max_bricks && nbricks > prm->max_bricks)
        return fail(rkf_set_error(RKFM_ERR_OOM, "sparse_build: %llu bricks > cap %llu", (unsigned long long)nbricks,
                                  (unsigned long long)prm->max_bricks));
    uint64_t need = ncell * 8 + (uint64_t)nblocks * 2048 + nbricks * 1536;
    if (c->memory_budget && used_bytes(c) + need > c->memory_budget)
        return fail(rkf_set_error(RKFM_ERR_OOM, "sparse_build: %llu bytes over the context budget",
                                  (unsigned long long)need));
    rc = s->l2.reserve(ncell * 4, false);
    if (!rc) rc = s->l1.reserve((uint64_t)nblocks * 2048 + 4, false);
    if (!rc) rc = s->dist.reserve(nbricks * 1024 + 1024, false);  // the first brick is read by unused loads
    if (!rc) rc = s->mat.reserve(nbricks * 512 + 4, false);
    if (!rc) rc = put(c->stream, *d_blk, blk_of.data(), (uint64_t)ncand * 4);
    if (!rc) rc = put(c->stream, *d_base, base.data(), (uint64_t)ncand * 4);
    if (!rc) rc = d_scand->reserve(nbricks * 4 + 4, false);
    if (!rc) rc = d_sbrick->reserve(nbricks * 4 + 4, false);
    if (rc) return fail(rc);
    s->l2.size = ncell * 4;
    s->l1.size = (uint64_t)nblocks * 2048;
    s->dist.size = nbricks * 1024;
    s->mat.size = nbricks * 512;
    rc = rkf_hip_check(hipMemsetAsync(s->l2.ptr, 0xFF, ncell * 4, c->stream), "l2 clear");
    if (rc) return fail(rc);
    if (ncand) {
        hipLaunchKernelGGL(rkf::assign_kernel, rkf::item_grid(ncand), dim3(512), 0, c->stream, B, ncand,
                           (const uint8_t *)d_flags->ptr,
                           (const uint32_t *)d_blk->ptr, (const uint32_t *)d_base->ptr, (uint32_t *)s->l2.ptr,
                           (uint32_t *)s->l1.ptr, (uint32_t *)d_scand->ptr, (uint32_t *)d_sbrick->ptr);
        rc = rkf_hip_check(hipGetLastError(), "assign launch");
        if (rc) return fail(rc);
    }
    if (nbricks) {
        hipLaunchKernelGGL(rkf::fill_kernel, rkf::item_grid(nbricks), dim3(512), 0, c->stream, B, nbricks,
                           (const uint32_t *)d_scand->ptr, (const uint32_t *)d_sbrick->ptr, (uint16_t *)s->dist.ptr,
                           (uint8_t *)s->mat.ptr);
        rc = rkf_hip_check(hipGetLastError(), "fill launch");
        if (rc) return fail(rc);
    }
    rc = rkf_hip_check(hipStreamSynchronize(c->stream), "sparse_build sync");
    if (rc) return fail(rc);
    s->info.n_blocks = nblocks;
    s->info.n_bricks = nbricks;
    rc = build_free_masks(c, s, block_cells);
    if (rc) return fail(rc);
    finish_info(s);
    c->scenes.push_back(s);
    *out = s;
    return RKFM_OK;
}

int rkfm_sparse_from_dense(rkfm_ctx *c, const uint32_t *map, const uint32_t dims[3], const rkf_voxel_sample *bricks,
                           uint64_t n_bricks, float vs, const float origin[3], rkfm_sparse_scene **out) {
    if (!c || !map || !dims || !origin || !out || (!bricks && n_bricks))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_from_dense: null argument");
    *out = nullptr;
    if (!(vs > 0.0f) || !dims[0] || !dims[1] || !dims[2])
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_from_dense: voxel size / dims");
    if (dims[0] > 8u * 32768u || dims[1] > 8u * 32768u || dims[2] > 8u * 32768u)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_from_dense: grid too large");
    uint32_t cells[3] = {(dims[0] + 7u) / 8u, (dims[1] + 7u) / 8u, (dims[2] + 7u) / 8u};
    uint64_t ncell = (uint64_t)cells[0] * cells[1] * cells[2];
    std::vector<uint32_t> l2(ncell, RKF_EMPTY_SLOT), l1, block_cells;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                uint32_t e = map[bx + (uint64_t)by * dims[0] + (uint64_t)bz * dims[0] * dims[1]];
                if (e == RKF_EMPTY_SLOT) continue;
                if (e != RKF_INTERIOR_SLOT && e >= n_bricks)
                    return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "sparse_from_dense: slot %u >= %llu bricks", e,
                                         (unsigned long long)n_bricks);
                uint64_t cell = (bx >> 3) + (uint64_t)(by >> 3) * cells[0] + (uint64_t)(bz >> 3) * cells[0] * cells[1];
                if (l2[cell] == RKF_EMPTY_SLOT) {
                    l2[cell] = (uint32_t)(l1.size() / 512);
                    l1.resize(l1.size() + 512, RKF_EMPTY_SLOT);
                    block_cells.push_back((uint32_t)cell);
                }
                l1[(uint64_t)l2[cell] * 512 + (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u] = e;
            }
    std::vector<uint16_t> dist(n_bricks * 512);
    std::vector<uint8_t> mat(n_bricks * 512);
    std::vector<uint16_t> aux(n_bricks * 512);
    bool has_aux = false;
    for (uint64_t i = 0; i < n_bricks * 512; i++) {
        uint32_t w0 = bricks[i].word0, w1 = bricks[i].word1;
        dist[i] = (uint16_t)(w0 & 0xFFFFu);
        mat[i] = (uint8_t)((w0 >> 16) & 0x3Fu);
        aux[i] = (uint16_t)(((w0 >> 22) & 0x3Fu) | ((w1 >> 24) << 8));
        has_aux |= aux[i] != 0;
    }
    rkfm_sparse_scene *s = new rkfm_sparse_scene();
    s->owner = c;
    for (int a = 0; a < 3; a++) {
        s->info.origin[a] = origin[a];
        s->info.cells[a] = cells[a];
    }
    s->info.voxel_size = vs;
    s->info.n_blocks = (uint32_t)(l1.size() / 512);
    s->info.n_bricks = n_bricks;
    int rc = put(c->stream, s->l2, l2.data(), ncell * 4);
    if (!rc) rc = put(c->stream, s->l1, l1.data(), l1.size() * 4);
    if (!rc) rc = put(c->stream, s->dist, dist.data(), dist.size() * 2);
    if (!rc) rc = put(c->stream, s->mat, mat.data(), mat.size());
    if (!rc && has_aux) rc = put(c->stream, s->aux, aux.data(), aux.size() * 2);
    if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "from_dense sync");
    if (!rc) rc = build_free_masks(c, s, block_cells);
    if (rc) {
        release_scene(s);
        delete s;
        return rc;
    }
    finish_info(s);
    c->scenes.push_back(s);
    *out = s;
    return RKFM_OK;
}

int rkfm_sparse_from_objects(rkfm_ctx *c, float vs, const float bmin[3], const float bmax[3],
                             rkfm_sparse_scene **out) {
    if (!c || !bmin || !bmax || !out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "from_objects: null argument");
    *out = nullptr;
    if (!(vs > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "from_objects: voxel size");
    const uint32_t nobj = c->su.num_objects;
    if (nobj > c->n_objects)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "from_objects: num_objects %u > %u uploaded", nobj, c->n_objects);
    if (nobj == 0) return rkf_set_error(RKFM_ERR_NOT_READY, "from_objects: no objects");
    const float bws = vs * 8.0f, cws = bws * 8.0f;
    uint32_t cells[3];
    for (int a = 0; a < 3; a++) {
        float ext = bmax[a] - bmin[a];
        if (!(ext > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "from_objects: empty bounds");
        double cc = std::ceil((double)ext / (double)cws);
        if (cc > 32768.0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "from_objects: grid too large");
        cells[a] = std::max(1u, (uint32_t)cc);
    }
    const uint64_t ncell = (uint64_t)cells[0] * cells[1] * cells[2];
    if (ncell > 0xFFFFFFF0ull) return rkf_set_error(RKFM_ERR_INVALID_ARG, "from_objects: too many cells");
    std::vector<rkf_gpu_object> objs(nobj);
    int rc = rkf_hip_check(hipMemcpy(objs.data(), c->objects.ptr, (uint64_t)nobj * sizeof(rkf_gpu_object),
                                     hipMemcpyDeviceToHost),
                           "from_objects: object records");
    if (rc) return rc;
    // Bin every object into the cells its AABB, widened by 3 bricks, meets.
    const double M = 3.0 * (double)bws;
    std::vector<uint64_t> pairs;
    for (uint32_t j = 0; j < nobj; j++) {
        if (objs[j].sdf_type == RKF_SDF_TYPE_NONE) continue;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)std::floor(((double)objs[j].aabb_min[a] - M - (double)bmin[a]) / (double)cws);
            hi[a] = (int64_t)std::floor(((double)objs[j].aabb_max[a] + M - (double)bmin[a]) / (double)cws);
            lo[a] = std::max<int64_t>(lo[a], 0);
            hi[a] = std::min<int64_t>(hi[a], (int64_t)cells[a] - 1);
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++)
                    pairs.push_back((((uint64_t)x + (uint64_t)y * cells[0] + (uint64_t)z * cells[0] * cells[1]) << 32) |
                                    j);
    }
    std::sort(pairs.begin(), pairs.end());
    std::vector<uint32_t> cand_cell, cand_off, cand_objs;
    for (uint64_t pr : pairs) {
        uint32_t cell = (uint32_t)(pr >> 32);
        if (cand_cell.empty() || cand_cell.back() != cell) {
            cand_cell.push_back(cell);
            cand_off.push_back((uint32_t)cand_objs.size());
        }
        cand_objs.push_back((uint32_t)(pr & 0xFFFFFFFFu));
    }
    cand_off.push_back((uint32_t)cand_objs.size());
    const uint32_t ncand = (uint32_t)cand_cell.size();

    rkfm_sparse_scene *s = new rkfm_sparse_scene();
    s->owner = c;
    for (int a = 0; a < 3; a++) {
        s->info.origin[a] = bmin[a];
        s->info.cells[a] = cells[a];
    }
    s->info.voxel_size = vs;
    TmpBufs tmp;
    DevBuf *d_cc = tmp.make(), *d_co = tmp.make(), *d_cg = tmp.make(), *d_flags = tmp.make(), *d_cnt = tmp.make();
    DevBuf *d_blk = tmp.make(), *d_base = tmp.make(), *d_scand = tmp.make(), *d_sbrick = tmp.make();
    DevBuf *d_any = tmp.make();
    auto fail = [&](int e) {
        release_scene(s);
        delete s;
        return e;
    };
    rc = put(c->stream, *d_cc, cand_cell.data(), cand_cell.size() * 4 + 4);
    if (!rc) rc = put(c->stream, *d_co, cand_off.data(), cand_off.size() * 4);
    if (!rc) rc = put(c->stream, *d_cg, cand_objs.data(), cand_objs.size() * 4 + 4);
    if (!rc) rc = d_flags->reserve((uint64_t)ncand * 512 + 4, false);
    if (!rc) rc = d_cnt->reserve((uint64_t)ncand * 4 + 4, false);
    if (!rc) rc = d_any->reserve(256, false);
    if (rc) return fail(rc);
    rkf::ObjBuildDev B;
    memset(&B, 0, sizeof(B));
    B.R.pool = (const rkf_voxel_sample *)c->pool.ptr;
    B.R.n_pool_voxels = c->n_pool_voxels;
    B.R.maps = (const uint32_t *)c->maps.ptr;
    B.R.n_maps = c->n_maps;
    B.R.objects = (const rkf_gpu_object *)c->objects.ptr;
    B.R.su = c->su;
    B.cand_cell = (const uint32_t *)d_cc->ptr;
    B.cand_off = (const uint32_t *)d_co->ptr;
    B.cand_objs = (const uint32_t *)d_cg->ptr;
    for (int a = 0; a < 3; a++) {
        B.origin[a] = bmin[a];
        B.cells[a] = cells[a];
    }
    B.vs = vs;
    std::vector<uint32_t> counts(ncand);
    if (ncand) {
        hipLaunchKernelGGL(rkf::obj_alloc_kernel, rkf::item_grid(ncand), dim3(512), 0, c->stream, B, ncand,
                           (uint8_t *)d_flags->ptr, (uint32_t *)d_cnt->ptr);
        rc = rkf_hip_check(hipGetLastError(), "obj_alloc launch");
        if (!rc)
            rc = rkf_hip_check(hipMemcpyAsync(counts.data(), d_cnt->ptr, (uint64_t)ncand * 4, hipMemcpyDeviceToHost,
                                              c->stream),
                               "counts download");
        if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "counts sync");
        if (rc) return fail(rc);
    }
    std::vector<uint32_t> blk_of(ncand, RKF_EMPTY_SLOT), base(ncand, 0), block_cells;
    uint64_t nbricks = 0;
    uint32_t nblocks = 0;
    for (uint32_t i = 0; i < ncand; i++) {
        if (!counts[i]) continue;
        blk_of[i] = nblocks++;
        block_cells.push_back(cand_cell[i]);
        base[i] = (uint32_t)nbricks;
        nbricks += counts[i];
    }
    if (nbricks >= 0xFFFFFFF0ull) return fail(rkf_set_error(RKFM_ERR_OOM, "from_objects: too many bricks"));
    uint64_t need = ncell * 8 + (uint64_t)nblocks * 2048 + nbricks * 2560;
    if (c->memory_budget && used_bytes(c) + need > c->memory_budget)
        return fail(rkf_set_error(RKFM_ERR_OOM, "from_objects: %llu bytes over the context budget",
                                  (unsigned long long)need));
    rc = s->l2.reserve(ncell * 4, false);
    if (!rc) rc = s->l1.reserve((uint64_t)nblocks * 2048 + 4, false);
    if (!rc) rc = s->dist.reserve(nbricks * 1024 + 1024, false);  // the first brick is read by unused loads
    if (!rc) rc = s->mat.reserve(nbricks * 512 + 4, false);
    if (!rc) rc = s->aux.reserve(nbricks * 1024 + 1024, false);
    if (!rc) rc = put(c->stream, *d_blk, blk_of.data(), (uint64_t)ncand * 4 + 4);
    if (!rc) rc = put(c->stream, *d_base, base.data(), (uint64_t)ncand * 4 + 4);
    if (!rc) rc = d_scand->reserve(nbricks * 4 + 4, false);
    if (!rc) rc = d_sbrick->reserve(nbricks * 4 + 4, false);
    if (rc) return fail(rc);
    s->l2.size = ncell * 4;
    s->l1.size = (uint64_t)nblocks * 2048;
    s->dist.size = nbricks * 1024;
    s->mat.size = nbricks * 512;
    s->aux.size = nbricks * 1024;
    rc = rkf_hip_check(hipMemsetAsync(s->l2.ptr, 0xFF, ncell * 4, c->stream), "l2 clear");
    if (!rc) rc = rkf_hip_check(hipMemsetAsync(d_any->ptr, 0, 4, c->stream), "aux flag clear");
    if (rc) return fail(rc);
    rkf::BuildDev A;
    memset(&A, 0, sizeof(A));
    A.cand_cell = B.cand_cell;
    if (ncand) {
        hipLaunchKernelGGL(rkf::assign_kernel, rkf::item_grid(ncand), dim3(512), 0, c->stream, A, ncand,
                           (const uint8_t *)d_flags->ptr, (const uint32_t *)d_blk->ptr, (const uint32_t *)d_base->ptr,
                           (uint32_t *)s->l2.ptr, (uint32_t *)s->l1.ptr, (uint32_t *)d_scand->ptr,
                           (uint32_t *)d_sbrick->ptr);
        rc = rkf_hip_check(hipGetLastError(), "assign launch");
        if (rc) return fail(rc);
    }
    uint32_t any_aux = 0;
    if (nbricks) {
        hipLaunchKernelGGL(rkf::obj_fill_kernel, rkf::item_grid(nbricks), dim3(512), 0, c->stream, B, nbricks,
                           (const uint32_t *)d_scand->ptr, (const uint32_t *)d_sbrick->ptr, (uint16_t *)s->dist.ptr,
                           (uint8_t *)s->mat.ptr, (uint16_t *)s->aux.ptr, (uint32_t *)d_any->ptr);
        rc = rkf_hip_check(hipGetLastError(), "obj_fill launch");
        if (!rc)
            rc = rkf_hip_check(hipMemcpyAsync(&any_aux, d_any->ptr, 4, hipMemcpyDeviceToHost, c->stream),
                               "aux flag download");
    }
    if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "from_objects sync");
    if (rc) return fail(rc);
    if (!any_aux) s->aux.release();  // no blend data: no aux plane, as for primitive scenes
    s->info.n_blocks = nblocks;
    s->info.n_bricks = nbricks;
    rc = build_free_masks(c, s, block_cells);
    if (rc) return fail(rc);
    finish_info(s);
    c->scenes.push_back(s);
    *out = s;
    return RKFM_OK;
}

int rkfm_sparse_download_aux(rkfm_ctx *c, const rkfm_sparse_scene *s, uint16_t *aux, uint32_t *present) {
    if (!c || !s || !present) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_download_aux: null argument");
    *present = s->aux.ptr ? 1u : 0u;
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "download sync");
    if (!rc && aux && s->aux.ptr && s->info.n_bricks)
        rc = rkf_hip_check(hipMemcpy(aux, s->aux.ptr, s->info.n_bricks * 1024, hipMemcpyDeviceToHost), "download aux");
    return rc;
}

int rkfm_sparse_info_get(const rkfm_sparse_scene *s, rkfm_sparse_info *info) {
    if (!s || !info) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_info: null argument");
    *info = s->info;
    return RKFM_OK;
}

int rkfm_sparse_destroy(rkfm_ctx *c, rkfm_sparse_scene *s) {
    if (!c || !s) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_destroy: null argument");
    auto it = std::find(c->scenes.begin(), c->scenes.end(), s);
    if (it == c->scenes.end()) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_destroy: scene not owned by ctx");
    (void)hipStreamSynchronize(c->stream);
    c->scenes.erase(it);
    release_scene(s);
    delete s;
    return RKFM_OK;
}

int rkfm_sparse_download(rkfm_ctx *c, const rkfm_sparse_scene *s, uint32_t *l2, uint32_t *l1, uint16_t *dist,
                         uint8_t *mat) {
    if (!c || !s) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_download: null argument");
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "download sync");
    if (!rc && l2) rc = rkf_hip_check(hipMemcpy(l2, s->l2.ptr, s->l2.size, hipMemcpyDeviceToHost), "download l2");
    if (!rc && l1 && s->l1.size) rc = rkf_hip_check(hipMemcpy(l1, s->l1.ptr, s->l1.size, hipMemcpyDeviceToHost), "download l1");
    if (!rc && dist && s->dist.size)
        rc = rkf_hip_check(hipMemcpy(dist, s->dist.ptr, s->dist.size, hipMemcpyDeviceToHost), "download dist");
    if (!rc && mat && s->mat.size) rc = rkf_hip_check(hipMemcpy(mat, s->mat.ptr, s->mat.size, hipMemcpyDeviceToHost), "download mat");
    return rc;
}

int rkfm_sparse_download_free(rkfm_ctx *c, const rkfm_sparse_scene *s, uint32_t *free_idx, uint64_t *free_bits,
                              uint32_t *oct, uint8_t *brick_ext, uint64_t *n_rows) {
    if (!c || !s || !n_rows) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_download_free: null argument");
    *n_rows = s->n_free_rows;
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "download sync");
    if (!rc && free_idx)
        rc = rkf_hip_check(hipMemcpy(free_idx, s->free_idx.ptr, s->free_idx.size, hipMemcpyDeviceToHost), "free idx");
    if (!rc && free_bits && s->free_bits.size)
        rc = rkf_hip_check(hipMemcpy(free_bits, s->free_bits.ptr, s->free_bits.size, hipMemcpyDeviceToHost), "free bits");
    if (!rc && oct) rc = rkf_hip_check(hipMemcpy(oct, s->oct.ptr, s->oct.size, hipMemcpyDeviceToHost), "octants");
    if (!rc && brick_ext && s->roww.size) {
        // octant planes (row * 8 + o) * 256 + lb / 2, low nibble for even lb
        std::vector<uint64_t> w(s->n_free_rows * 512u);
        rc = rkf_hip_check(hipMemcpy(w.data(), s->roww.ptr, w.size() * 8, hipMemcpyDeviceToHost), "brick extents");
        for (uint64_t r = 0; !rc && r < s->n_free_rows; r++)
            for (uint32_t o = 0; o < 8u; o++)
                for (uint32_t j = 0; j < 256u; j++) {
                    const uint32_t lo = (uint32_t)(w[r * 512u + 2u * j] >> (32u + 4u * o)) & 15u;
                    const uint32_t hi = (uint32_t)(w[r * 512u + 2u * j + 1u] >> (32u + 4u * o)) & 15u;
                    brick_ext[(r * 8u + o) * 256u + j] = (uint8_t)(lo | (hi << 4));
                }
    }
    return rc;
}

// Tile intervals into `tiles` (the context's large-rectangle list as
// scratch; stream-ordered, so successive calls on one stream may share it).
int rkf_tile_camera_check(const rkf_camera_uniforms *cam) {
    if (!(fabsf(cam->jitter[0]) <= 0.5f && fabsf(cam->jitter[1]) <= 0.5f))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile intervals: |jitter| above 0.5");
    {
        // the kernels invert [forward right up]; refuse a degenerate basis
        double f[3], r[3], u[3];
        for (int a = 0; a < 3; a++) {
            f[a] = cam->forward[a];
            r[a] = cam->right[a];
            u[a] = cam->up[a];
        }
        const double det = f[0] * (r[1] * u[2] - r[2] * u[1]) + f[1] * (r[2] * u[0] - r[0] * u[2]) +
                           f[2] * (r[0] * u[1] - r[1] * u[0]);
        if (!std::isfinite(det) || fabs(det) < 1e-12)
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile intervals: degenerate camera basis");
    }
    return RKFM_OK;
}

constexpr uint64_t kOrderMaxTiles = 65536u;

// clear = false: tiles, cnt (with its sort words) and the big-rectangle
// count are known to be cleared already.
static int tile_intervals(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                          const rkf_scene_uniforms *su, uint32_t w, uint32_t h, uint2 *tiles, uint32_t *cnt,
                          bool clear, hipStream_t st) {
    int rc = rkf_tile_camera_check(cam);
    const void *count0 = c->tile_big_count.ptr;
    if (!rc) rc = c->tile_big.reserve((uint64_t)s->info.n_blocks * sizeof(rkf::TileBig) + 64, false);
    if (!rc) rc = c->tile_big_count.reserve(256, false);
    if (rc) return rc;
    if (c->tile_big_count.ptr != count0) clear = true;  // a fresh allocation holds no cleared count
    rkf::SparseDev D = s->dev();
    return rkf_hip_check(rkf_launch_sparse_tiles(&D, cam, su->max_distance, w, h, tiles, cnt,
                                                 (rkf::TileBig *)c->tile_big.ptr,
                                                 (uint32_t *)c->tile_big_count.ptr, clear, st),
                         "tile intervals launch");
}

int rkfm_sparse_tile_intervals(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                               const rkf_scene_uniforms *su, uint32_t w, uint32_t h, float *tiles, void *stream) {
    if (!c || !s || !cam || !su || !tiles) return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile intervals: null argument");
    if (w == 0 || h == 0 || w > 32768u || h > 32768u)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile intervals: %ux%u", w, h);
    c->iv_clean = 0;  // the big-rectangle count is shared with the context's views
    return tile_intervals(c, s, cam, su, w, h, (uint2 *)tiles, nullptr, true, stream ? (hipStream_t)stream : c->stream);
}

// The tile counts and launch order of a culled view into the context's
// buffers (tile_iv, tile_cnt with its 512 sort words, tile_order).
static int tile_order(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                      const rkf_scene_uniforms *su, uint32_t w, uint32_t h, hipStream_t st) {
    const uint64_t n_tiles = (uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u);
    const void *iv0 = c->tile_iv.ptr, *cnt0 = c->tile_cnt.ptr;
    int rc = c->tile_iv.reserve(n_tiles * 8u, false);
    // 512 bin words of the order's counting sort, then one count per tile
    if (!rc) rc = c->tile_cnt.reserve((n_tiles + 512u) * 4u, false);
    if (!rc) rc = c->tile_order.reserve(n_tiles * 4u, false);
    const bool clean = c->tile_iv.ptr == iv0 && c->tile_cnt.ptr == cnt0 && c->clean_stream == st &&
                       c->iv_clean >= n_tiles && c->cnt_clean >= n_tiles;
    c->iv_clean = c->cnt_clean = 0;
    if (!rc)
        rc = tile_intervals(c, s, cam, su, w, h, (uint2 *)c->tile_iv.ptr, (uint32_t *)c->tile_cnt.ptr + 512u, !clean,
                            st);
    if (!rc)
        rc = rkf_hip_check(rkf_launch_sparse_tile_order((uint32_t *)c->tile_cnt.ptr + 512u, (uint32_t)n_tiles,
                                                        (uint32_t *)c->tile_order.ptr, st),
                           "tile order launch");
    return rc;
}

int rkfm_sparse_tile_order(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                           const rkf_scene_uniforms *su, uint32_t w, uint32_t h, uint32_t *counts, uint32_t *order,
                           void *stream) {
    if (!c || !s || !cam || !su || !counts || !order)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile order: null argument");
    if (w == 0 || h == 0 || w > 32768u || h > 32768u) return rkf_set_error(RKFM_ERR_INVALID_ARG, "tile order: %ux%u", w, h);
    const uint64_t n_tiles = (uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u);
    if (n_tiles > kOrderMaxTiles)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "tile order: %llu tiles (max %llu)", (unsigned long long)n_tiles,
                             (unsigned long long)kOrderMaxTiles);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = tile_order(c, s, cam, su, w, h, st);
    if (!rc)
        rc = rkf_hip_check(hipMemcpyAsync(counts, (uint32_t *)c->tile_cnt.ptr + 512u, n_tiles * 4u,
                                          hipMemcpyDeviceToDevice, st),
                           "tile order counts copy");
    if (!rc)
        rc = rkf_hip_check(hipMemcpyAsync(order, c->tile_order.ptr, n_tiles * 4u, hipMemcpyDeviceToDevice, st),
                           "tile order copy");
    return rc;
}

int rkfm_sparse_march(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                      const rkf_scene_uniforms *su, const rkfm_scene_ext *ext, const rkfm_gbuffer *out,
                      uint32_t variant, void *stream) {
    if (!c || !s || !cam || !su) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_march: null argument");
    if (variant > 2) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_march: variant %u", variant);
    if (ext && ext->hit_eps_mode > RKFM_EPS_RELATIVE)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_march: hit_eps_mode %u", ext->hit_eps_mode);
    rkf::SparseMarchParams P;
    memset(&P, 0, sizeof(P));
    int rc = rkf_fill_gbuf(out, P.g);
    if (rc) return rc;
    P.S = s->dev();
    P.cam = *cam;
    P.su = *su;
    if (ext) P.ext = *ext;
    P.object_id = 0;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (ext && (ext->march_flags & ~RKFM_MARCH_TILE_CULL))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_march: march_flags 0x%x", ext->march_flags);
    if (ext && (ext->march_flags & RKFM_MARCH_TILE_CULL)) {
        const uint32_t tx = (out->width + 7u) / 8u, ty = (out->height + 7u) / 8u;
        const uint64_t n_tiles = (uint64_t)tx * ty;
        // The plain kernel launches the tiles with most occupied cells first
        // when the frame is a few waves deep per resident wave slot (its last
        // long tiles then set the launch's end: 1080p, 32,400 tiles, 0.291 ->
        // 0.270 ms on MI355X); a deeper frame keeps row-major order, which
        // keeps neighbouring tiles' bricks in L2 together (2160p, 129,600
        // tiles: 1.69 ms row-major, 1.73 ms ordered).
        const bool order = variant == 0 && n_tiles <= kOrderMaxTiles;
        if (order) {
            rc = tile_order(c, s, cam, su, out->width, out->height, st);
        } else {
            const void *iv0 = c->tile_iv.ptr;
            rc = c->tile_iv.reserve(n_tiles * 8u, false);
            const bool clean = c->tile_iv.ptr == iv0 && c->clean_stream == st && c->iv_clean >= n_tiles;
            c->iv_clean = 0;
            if (!rc)
                rc = tile_intervals(c, s, cam, su, out->width, out->height, (uint2 *)c->tile_iv.ptr, nullptr, !clean,
                                    st);
        }
        if (rc) return rc;
        P.tiles = (const uint2 *)c->tile_iv.ptr;
        P.tiles_x = tx;
        if (order) P.order = (const uint32_t *)c->tile_order.ptr;
        if (variant == 0) {  // the plain march puts what it read back to the cleared state
            P.reset_iv = (uint2 *)c->tile_iv.ptr;
            P.reset_big = (uint32_t *)c->tile_big_count.ptr;
            if (order) {
                P.reset_cnt = (uint32_t *)c->tile_cnt.ptr + 512u;
                P.reset_bins = (uint32_t *)c->tile_cnt.ptr;
            }
        }
    }
    if (variant == 1 && !c->tuned_waves) {
        rc = rkf_hip_check(rkf_occupancy_waves(0, &c->tuned_waves), "occupancy");
        if (rc) return rc;
    }
    if (variant == 2) {
        uint64_t px = (uint64_t)out->width * out->height;
        rc = c->march_hits.reserve(px * rkf_hit_record_bytes(), false);
        if (!rc) rc = c->march_count.reserve(256, false);
        if (rc) return rc;
        uint32_t shade_waves = 0;
        rc = rkf_device_waves(16, shade_waves);
        if (rc) return rc;
        return rkf_hip_check(rkf_launch_sparse_march_split(&P, c->march_hits.ptr, (uint32_t *)c->march_count.ptr,
                                                           shade_waves, st),
                             "sparse_march launch");
    }
    rc = rkf_hip_check(rkf_launch_sparse_march(&P, variant, c->tuned_waves, st), "sparse_march launch");
    if (!rc && P.reset_iv) {  // cleared again once this launch is done, on st
        const uint64_t n_tiles = (uint64_t)P.tiles_x * ((out->height + 7u) / 8u);
        c->iv_clean = n_tiles;
        if (P.reset_cnt) c->cnt_clean = n_tiles;
        c->clean_stream = st;
    }
    return rc;
}

int rkfm_sparse_contacts(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkfm_body *bodies, uint32_t n,
                         rkfm_contact *contacts, uint32_t *counts, void *stream) {
    if (!c || !s || (n && (!bodies || !contacts || !counts)))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "contacts: null argument");
    if (n > (1u << 27)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "contacts: %u bodies", n);
    rkf::SparseDev D = s->dev();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    return rkf_hip_check(rkf_launch_sparse_contacts(&D, bodies, n, contacts, counts, st), "contacts launch");
}

int rkfm_sparse_sphere_cast(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkfm_sphere_cast *casts, uint32_t n,
                            rkfm_cast_result *out, void *stream) {
    if (!c || !s || (n && (!casts || !out))) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sphere_cast: null argument");
    rkf::SparseDev D = s->dev();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    return rkf_hip_check(rkf_launch_sparse_cast(&D, casts, n, out, st), "sphere cast launch");
}

int rkfm_sparse_query_points(rkfm_ctx *c, const rkfm_sparse_scene *s, const float *xyz, uint32_t n,
                             rkfm_point_result *out, void *stream) {
    if (!c || !s || (n && (!xyz || !out))) return rkf_set_error(RKFM_ERR_INVALID_ARG, "query_points: null argument");
    rkf::SparseDev D = s->dev();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    return rkf_hip_check(rkf_launch_sparse_query(&D, xyz, n, out, st), "query launch");
}

}  // extern "C"
