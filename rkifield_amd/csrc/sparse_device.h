// sparse_device.h — per-ray device code for resident sparse scenes (two-level
// index + SoA pool), shared by the single-view kernels and the batched launch.
//
// Field: the reference's sample_voxelized (ray_march.wgsl:353-413) for one
// world-space object whose dense brick map is the expansion of the two-level
// index, with grid coordinates multiplied by inv_vs (1/vs rounded to float)
// where the WGSL divides by vs (its f32 division is specified to 2.5 ULP;
// the multiply is one instruction, a correctly rounded division a dozen):
// 8 corner voxels, EMPTY -> +8 vs, INTERIOR -> -2 vs (stepping) or
// -8 vs (gradient, :300-315), trilinear mix x->y->z, plus the outside-grid
// distance. Step rules: ray_march_tiled (:905-982): near-step budget
// (d < 0.02), MIN_STEP, 8-step bisection on overstep, hit threshold
// (absolute or relative). Normals: 6-tap central difference at 2 vs
// (:504-513, :1003-1020). G-buffer: :1083-1111. Shadows: soft_shadow
// (shade_common_shading.wgsl:9-33) from shade_pbr.wgsl:40's origin.
//
// Multi-level free-space DDA (no reference counterpart): with
// g = (p - origin) * inv_vs - 0.5 and b0 = floor(floor(g)/8), a sample's corners lie
// in bricks {b0, b0+1}^3. If all of them are EMPTY the field is the constant
// 8 vs over the region floor(g/8) == b0, and over the cube of free bricks
// that has b0 at its corner and extends E_b bricks in the directions the ray
// travels (E_b = b0's brick extent for the ray's octant, at most 15). When
// the whole L2 cell is free, the region is the cube of free cells that has
// this cell at its corner and extends E cells in those directions (E = the
// cell's extent for the ray's octant, at most 15). The ray jumps to the region's exit t_x,
// takes prev_t = t_x, prev_dist = 8 vs and resumes at t_x + vs/256. One skip
// counts as one loop iteration.
//
// Tile intervals (optional): a view's 8x8-pixel tiles carry [t_lo, t_hi],
// the range of ray parameters at which any of the tile's rays can come
// within 9 vs (25 vs on grid-boundary sides) of an L2 cell holding an L1
// block; outside it every sample is the free value (a sample reads bricks
// {b0, b0+1}^3, i.e. points 7.5 vs below to 0.5 vs above a brick, and
// samples clamped at the grid boundary reach 16 vs further).
#pragma once

#include "kernels.h"

namespace rkf {

struct SpCtx {
    const SparseDev *S;
    uint32_t evals;
};


RKF_DEV uint32_t sp_slot(const SparseDev &S, int32_t cx, int32_t cy, int32_t cz) {
    uint32_t bx = (uint32_t)cx >> 3, by = (uint32_t)cy >> 3, bz = (uint32_t)cz >> 3;
    uint64_t cell = (uint64_t)(bx >> 3) + (uint64_t)(by >> 3) * S.cells[0] +
                    (uint64_t)(bz >> 3) * S.cells[0] * S.cells[1];
    uint32_t blk = S.l2[cell];
    if (blk == RKF_EMPTY_SLOT) return RKF_EMPTY_SLOT;
    return S.l1[(uint64_t)blk * 512u + (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u];
}

RKF_DEV float sp_voxel(const SparseDev &S, int32_t cx, int32_t cy, int32_t cz, float vs, bool grad) {
    cx = clampi(cx, 0, S.tot[0] - 1);
    cy = clampi(cy, 0, S.tot[1] - 1);
    cz = clampi(cz, 0, S.tot[2] - 1);
    uint32_t slot = sp_slot(S, cx, cy, cz);
    if (slot == RKF_EMPTY_SLOT) return vs * 8.0f;
    if (slot == RKF_INTERIOR_SLOT) return grad ? -(vs * 8.0f) : -(vs * 2.0f);
    uint64_t idx = (uint64_t)slot * 512u + (uint32_t)(cx & 7) + (uint32_t)(cy & 7) * 8u + (uint32_t)(cz & 7) * 64u;
    return h2f(S.dist[idx]);
}

// pre: the slot of the brick holding corner (x0, y0, z0) when the caller
// already loaded it (sp_free), else kNoSlot.
RKF_DEV float sp_sample(SpCtx &c, f3 p, bool grad, bool count, uint32_t pre = kNoSlot) {
    const SparseDev &S = *c.S;
    float vs = S.vs;
    float be = vs * 8.0f;
    f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    f3 gp = sub(p, mk(S.origin[0], S.origin[1], S.origin[2]));
    float lo = vs * 0.01f;
    f3 cl = mk(clampf(gp.x, lo, gs.x - lo), clampf(gp.y, lo, gs.y - lo), clampf(gp.z, lo, gs.z - lo));
    float outside = length(sub(gp, cl));
    if (outside > be * 2.0f) return outside;
    f3 vc = mk(cl.x * S.inv_vs - 0.5f, cl.y * S.inv_vs - 0.5f, cl.z * S.inv_vs - 0.5f);
    int32_t x0 = (int32_t)floorf(vc.x), y0 = (int32_t)floorf(vc.y), z0 = (int32_t)floorf(vc.z);
    float tx = vc.x - (float)x0, ty = vc.y - (float)y0, tz = vc.z - (float)z0;
    if (count) c.evals++;
    float c000, c100, c010, c110, c001, c101, c011, c111;
    // Fast path: all 8 corners inside one brick (no clamping, no brick
    // crossing) -> one index walk and 8 loads from one 1 KiB distance page.
    bool one_brick = x0 >= 0 && y0 >= 0 && z0 >= 0 && x0 + 1 < S.tot[0] && y0 + 1 < S.tot[1] &&
                     z0 + 1 < S.tot[2] && (x0 & 7) != 7 && (y0 & 7) != 7 && (z0 & 7) != 7;
    if (one_brick) {
        uint32_t slot = pre != kNoSlot ? pre : sp_slot(S, x0, y0, z0);
        if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) {
            float v = slot == RKF_EMPTY_SLOT ? vs * 8.0f : (grad ? -(vs * 8.0f) : -(vs * 2.0f));
            c000 = c100 = c010 = c110 = c001 = c101 = c011 = c111 = v;
        } else {
            const uint16_t *pg = S.dist + (uint64_t)slot * 512u + (uint32_t)(x0 & 7) + (uint32_t)(y0 & 7) * 8u +
                                 (uint32_t)(z0 & 7) * 64u;
            c000 = h2f(pg[0]);
            c100 = h2f(pg[1]);
            c010 = h2f(pg[8]);
            c110 = h2f(pg[9]);
            c001 = h2f(pg[64]);
            c101 = h2f(pg[65]);
            c011 = h2f(pg[72]);
            c111 = h2f(pg[73]);
        }
    } else {
        c000 = sp_voxel(S, x0, y0, z0, vs, grad);
        c100 = sp_voxel(S, x0 + 1, y0, z0, vs, grad);
        c010 = sp_voxel(S, x0, y0 + 1, z0, vs, grad);
        c110 = sp_voxel(S, x0 + 1, y0 + 1, z0, vs, grad);
        c001 = sp_voxel(S, x0, y0, z0 + 1, vs, grad);
        c101 = sp_voxel(S, x0 + 1, y0, z0 + 1, vs, grad);
        c011 = sp_voxel(S, x0, y0 + 1, z0 + 1, vs, grad);
        c111 = sp_voxel(S, x0 + 1, y0 + 1, z0 + 1, vs, grad);
    }
    float c00 = mix(c000, c100, tx), c10 = mix(c010, c110, tx);
    float c01 = mix(c001, c101, tx), c11 = mix(c011, c111, tx);
    float c0 = mix(c00, c10, ty), c1 = mix(c01, c11, ty);
    return mix(c0, c1, tz) + outside;
}

// Slots of the (at most 2 x 2 x 2) bricks holding every corner voxel of the
// samples within `reach` voxels of p on each axis, looked up once so that the
// shading samples around a hit (6 gradient taps at +-2 vs, the value at the
// hit) share one index walk. sp_sample_bc falls back to the index for a corner
// outside the cached bricks, so results equal sp_voxel's exactly.
struct BrickCache {
    int32_t b0[3];
    uint32_t span[3];  // 1 or 2 bricks cached per axis
    uint32_t slot[8];  // brick (b0 + (k & 1), b0 + (k >> 1 & 1), b0 + (k >> 2))
};

RKF_DEV uint32_t sp_brick_slot(const SparseDev &S, uint32_t bx, uint32_t by, uint32_t bz) {
    uint64_t cell = (uint64_t)(bx >> 3) + (uint64_t)(by >> 3) * S.cells[0] +
                    (uint64_t)(bz >> 3) * S.cells[0] * S.cells[1];
    uint32_t blk = S.l2[cell];
    if (blk == RKF_EMPTY_SLOT) return RKF_EMPTY_SLOT;
    return S.l1[(uint64_t)blk * 512u + (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u];
}

RKF_DEV void bc_init(const SparseDev &S, f3 p, int32_t reach, BrickCache &bc) {
    f3 gp = sub(p, mk(S.origin[0], S.origin[1], S.origin[2]));
    float g[3] = {gp.x * S.inv_vs - 0.5f, gp.y * S.inv_vs - 0.5f, gp.z * S.inv_vs - 0.5f};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // clampf keeps the float -> int conversion in range for far points
        float f = floorf(clampf(g[a], -64.0f, (float)S.tot[a] + 64.0f));
        int32_t lo = clampi((int32_t)f - reach, 0, S.tot[a] - 1);
        int32_t hi = clampi((int32_t)f + 1 + reach, 0, S.tot[a] - 1);
        bc.b0[a] = lo >> 3;
        bc.span[a] = (uint32_t)((hi >> 3) - (lo >> 3)) + 1u;
    }
    // sp_brick_slot for each cached brick: the L2 entries in one round, the
    // L1 slots in a second (unused loads read entry 0)
    uint32_t blk[8], lb[8];
    bool need[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        const uint32_t dx = k & 1u, dy = (k >> 1) & 1u, dz = k >> 2;
        need[k] = dx < bc.span[0] && dy < bc.span[1] && dz < bc.span[2];
        const uint32_t bx = (uint32_t)bc.b0[0] + dx, by = (uint32_t)bc.b0[1] + dy, bz = (uint32_t)bc.b0[2] + dz;
        const uint64_t cell = (uint64_t)(bx >> 3) + (uint64_t)(by >> 3) * S.cells[0] +
                              (uint64_t)(bz >> 3) * S.cells[0] * S.cells[1];
        lb[k] = (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u;
        blk[k] = S.l2[need[k] ? cell : 0u];
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        loaded(blk[k]);
        need[k] = need[k] && blk[k] != RKF_EMPTY_SLOT;
        bc.slot[k] = S.l1[need[k] ? (uint64_t)blk[k] * 512u + lb[k] : 0u];
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        loaded(bc.slot[k]);
        bc.slot[k] = need[k] ? bc.slot[k] : RKF_EMPTY_SLOT;
    }
}

// Slot of the brick holding voxel (cx, cy, cz) (clamped to the grid): from
// the cache by selects when the brick is cached (in_cache), else not set.
RKF_DEV uint32_t bc_slot(const BrickCache &bc, int32_t cx, int32_t cy, int32_t cz, bool &in_cache) {
    const uint32_t dx = (uint32_t)((cx >> 3) - bc.b0[0]), dy = (uint32_t)((cy >> 3) - bc.b0[1]),
                   dz = (uint32_t)((cz >> 3) - bc.b0[2]);
    in_cache = dx < bc.span[0] && dy < bc.span[1] && dz < bc.span[2];
    const uint32_t *s = bc.slot;
    const bool ox = dx != 0u, oy = dy != 0u, oz = dz != 0u;
    const uint32_t a0 = ox ? s[1] : s[0], a1 = ox ? s[3] : s[2], a2 = ox ? s[5] : s[4], a3 = ox ? s[7] : s[6];
    const uint32_t b0 = oy ? a1 : a0, b1 = oy ? a3 : a2;
    return oz ? b1 : b0;
}

// sp_sample with corners from a brick cache (same arithmetic, same results):
// the 8 corner slots by selects (the index for a corner outside the cached
// bricks, only when some lane has one), then the 8 distance loads in one
// round.
RKF_DEV float sp_sample_bc(const SparseDev &S, const BrickCache &bc, f3 p, bool grad) {
    float vs = S.vs;
    float be = vs * 8.0f;
    f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    f3 gp = sub(p, mk(S.origin[0], S.origin[1], S.origin[2]));
    float lo = vs * 0.01f;
    f3 cl = mk(clampf(gp.x, lo, gs.x - lo), clampf(gp.y, lo, gs.y - lo), clampf(gp.z, lo, gs.z - lo));
    float outside = length(sub(gp, cl));
    if (outside > be * 2.0f) return outside;
    f3 vc = mk(cl.x * S.inv_vs - 0.5f, cl.y * S.inv_vs - 0.5f, cl.z * S.inv_vs - 0.5f);
    int32_t x0 = (int32_t)floorf(vc.x), y0 = (int32_t)floorf(vc.y), z0 = (int32_t)floorf(vc.z);
    float tx = vc.x - (float)x0, ty = vc.y - (float)y0, tz = vc.z - (float)z0;
    const int32_t X[2] = {clampi(x0, 0, S.tot[0] - 1), clampi(x0 + 1, 0, S.tot[0] - 1)};
    const int32_t Y[2] = {clampi(y0, 0, S.tot[1] - 1), clampi(y0 + 1, 0, S.tot[1] - 1)};
    const int32_t Z[2] = {clampi(z0, 0, S.tot[2] - 1), clampi(z0 + 1, 0, S.tot[2] - 1)};
    uint32_t sl[8];
    bool inc[8];
    bool miss = false;
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        sl[k] = bc_slot(bc, X[k & 1u], Y[(k >> 1) & 1u], Z[k >> 2], inc[k]);
        miss = miss || !inc[k];
    }
    if (__any(miss)) {
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++)
            if (!inc[k]) sl[k] = sp_slot(S, X[k & 1u], Y[(k >> 1) & 1u], Z[k >> 2]);
    }
    uint32_t raw[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        const bool live = sl[k] < RKF_INTERIOR_SLOT;
        const uint32_t loc = (uint32_t)(X[k & 1u] & 7) + (uint32_t)(Y[(k >> 1) & 1u] & 7) * 8u +
                             (uint32_t)(Z[k >> 2] & 7) * 64u;
        raw[k] = S.dist[(uint64_t)(live ? sl[k] : 0u) * 512u + (live ? loc : 0u)];
    }
    const float v_empty = vs * 8.0f, v_interior = grad ? -(vs * 8.0f) : -(vs * 2.0f);
    float c[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        loaded(raw[k]);
        const uint32_t sent = __float_as_uint(sl[k] == RKF_EMPTY_SLOT ? v_empty : v_interior);
        c[k] = __uint_as_float(sl[k] < RKF_INTERIOR_SLOT ? __float_as_uint(h2f(raw[k])) : sent);
    }
    float c00 = mix(c[0], c[1], tx), c10 = mix(c[2], c[3], tx);
    float c01 = mix(c[4], c[5], tx), c11 = mix(c[6], c[7], tx);
    float c0 = mix(c00, c10, ty), c1 = mix(c01, c11, ty);
    return mix(c0, c1, tz) + outside;
}

// Nearest voxel floor(grid_pos * inv_vs) as ray_march.wgsl:540-596. Returns the
// pool index or ~0.
RKF_DEV uint64_t sp_nearest(const SparseDev &S, f3 p, uint32_t &slot_o, uint32_t &vox_o, uint32_t &cell_o,
                            const BrickCache &bc) {
    float vs = S.vs;
    float be = vs * 8.0f;
    f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    f3 gp = sub(p, mk(S.origin[0], S.origin[1], S.origin[2]));
    slot_o = RKF_EMPTY_SLOT;
    vox_o = 0xFFFFFFFFu;
    cell_o = 0xFFFFFFFFu;
    if (gp.x < 0.0f || gp.y < 0.0f || gp.z < 0.0f || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z) return ~0ull;
    int32_t cx = clampi((int32_t)floorf(gp.x * S.inv_vs), 0, S.tot[0] - 1);
    int32_t cy = clampi((int32_t)floorf(gp.y * S.inv_vs), 0, S.tot[1] - 1);
    int32_t cz = clampi((int32_t)floorf(gp.z * S.inv_vs), 0, S.tot[2] - 1);
    bool in_cache = false;
    uint32_t slot = bc_slot(bc, cx, cy, cz, in_cache);
    if (!in_cache) slot = sp_slot(S, cx, cy, cz);
    cell_o = ((uint32_t)cx >> 6) + ((uint32_t)cy >> 6) * S.cells[0] + ((uint32_t)cz >> 6) * S.cells[0] * S.cells[1];
    slot_o = slot;
    uint32_t lv = (uint32_t)(cx & 7) + (uint32_t)(cy & 7) * 8u + (uint32_t)(cz & 7) * 64u;
    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) return ~0ull;
    vox_o = lv;
    return (uint64_t)slot * 512u + lv;
}

RKF_DEV uint32_t sp_material(const SparseDev &S, f3 p, const BrickCache &bc) {
    uint32_t a, b, d;
    uint64_t i = sp_nearest(S, p, a, b, d, bc);
    return i == ~0ull ? 0u : (uint32_t)(S.mat[i] & 0x3Fu);
}

// Free-space test + exit parameter (see file header). One load of the cell
// word decides cell-level skips; otherwise one load of the brick's row word
// gives its extent and its L1 slot, so a step that must sample has its slot
// (returned in `slot`, else kNoSlot) after two dependent loads.
RKF_DEV bool sp_free(const SparseDev &S, f3 pos, f3 sd, f3 inv, float t, float &t_exit, uint32_t &slot) {
    slot = kNoSlot;
    float vs = S.vs;
    float be = vs * 8.0f;
    f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    f3 gp = sub(pos, mk(S.origin[0], S.origin[1], S.origin[2]));
    float lo = vs * 0.01f;
    if (gp.x < lo || gp.y < lo || gp.z < lo || gp.x > gs.x - lo || gp.y > gs.y - lo || gp.z > gs.z - lo) return false;
    float g[3] = {gp.x * S.inv_vs - 0.5f, gp.y * S.inv_vs - 0.5f, gp.z * S.inv_vs - 0.5f};
    int32_t v0[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        v0[a] = (int32_t)floorf(g[a]);
        if (v0[a] < 0 || v0[a] > S.tot[a] - 2) return false;
    }
    uint32_t b[3] = {(uint32_t)v0[0] >> 3, (uint32_t)v0[1] >> 3, (uint32_t)v0[2] >> 3};
    uint32_t q[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    uint64_t cell = (uint64_t)q[0] + (uint64_t)q[1] * S.cells[0] + (uint64_t)q[2] * S.cells[0] * S.cells[1];
    float lo_b[3], hi_b[3];
    const uint64_t cw = S.cellw[cell];
    const uint32_t octant = (sd.x < 0.0f ? 1u : 0u) | (sd.y < 0.0f ? 2u : 0u) | (sd.z < 0.0f ? 4u : 0u);
    const uint32_t e = ((uint32_t)cw >> (4u * octant)) & 15u;
    if (e > 0) {
        // the cube of cells q + [0, e) in the travel direction of each axis is free
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)q[a] - (float)(e - 1u)) * 64.0f;
            hi_b[a] = ((float)q[a] + (float)e) * 64.0f;
        }
    } else {
        uint32_t row = (uint32_t)(cw >> 32);
        uint32_t lb = (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u;
        const uint64_t rw = S.roww[(uint64_t)row * 512u + lb];
        const uint32_t eb = (uint32_t)(rw >> (32u + 4u * octant)) & 15u;
        if (eb == 0) {
            slot = (uint32_t)rw;
            return false;
        }
        // the cube of bricks b + [0, eb) in the travel direction of each axis is free
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)b[a] - (float)(eb - 1u)) * 8.0f;
            hi_b[a] = ((float)b[a] + (float)eb) * 8.0f;
        }
    }
    float dv[3] = {sd.x, sd.y, sd.z};
    float iv[3] = {inv.x, inv.y, inv.z};
    float dt = kMaxFloat;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float bound = dv[a] > 0.0f ? hi_b[a] : lo_b[a];
        float da = (bound - g[a]) * iv[a];
        dt = fminf(dt, da);
    }
    t_exit = t + fmaxf(dt, 0.0f);
    return true;
}

// The slots of the 2x2x2 bricks b0 + {0,1}^3 that a sample whose lower
// corner voxel lies in brick b0 reads, kept per ray while its samples stay in
// b0 (a brick that is not free): a step then needs only the distance loads.
struct BrickNbhd {
    uint64_t key;      // b0's brick coordinates, 21 bits each (grids hold at most 2^18
                       // bricks per axis); ~0 when empty
    uint32_t slot[8];  // brick b0 + (k & 1, k >> 1 & 1, k >> 2); constant indices only
    uint32_t cell;     // last cell met that is not free for this ray's octant, ~0 when none
    uint32_t row;      // its free-mask row
    uint32_t nrow;     // 8-lane groups: the row of neighbour cell (k & 1, k >> 1 & 1, k >> 2) of `cell`
};

RKF_DEV void nb_reset(BrickNbhd &n) {
    n.key = ~0ull;
    n.cell = ~0u;
}

// Slot of corner brick (ox, oy, oz) in {0,1}^3 by selects, so the slots stay
// in registers.
RKF_DEV uint32_t nb_pick(const BrickNbhd &n, bool ox, bool oy, bool oz) {
    const uint32_t *s = n.slot;
    const uint32_t a0 = ox ? s[1] : s[0], a1 = ox ? s[3] : s[2], a2 = ox ? s[5] : s[4], a3 = ox ? s[7] : s[6];
    const uint32_t b0 = oy ? a1 : a0, b1 = oy ? a3 : a2;
    return oz ? b1 : b0;
}

// sp_sample (stepping semantics) for an in-grid point whose lower corner
// voxel v0 (= floor(g)) lies in the cached brick: same corner values, same
// arithmetic (the outside-grid distance is +0 here). Branch-free per lane:
// every corner's load is issued (a sentinel slot reads the pool's first
// brick, the pool always holds one) and the sentinel values are selected
// afterwards.
RKF_DEV float nb_sample(const SparseDev &S, const BrickNbhd &n, const float g[3], const int32_t v0[3]) {
    const float vs = S.vs;
    const float v_empty = vs * 8.0f, v_interior = -(vs * 2.0f);
    const int32_t x0 = v0[0], y0 = v0[1], z0 = v0[2];
    const float tx = g[0] - (float)x0, ty = g[1] - (float)y0, tz = g[2] - (float)z0;
    const bool ox = (x0 & 7) == 7, oy = (y0 & 7) == 7, oz = (z0 & 7) == 7;
    const uint32_t lx[2] = {(uint32_t)(x0 & 7), (uint32_t)((x0 + 1) & 7)};
    const uint32_t ly[2] = {(uint32_t)(y0 & 7) * 8u, (uint32_t)((y0 + 1) & 7) * 8u};
    const uint32_t lz[2] = {(uint32_t)(z0 & 7) * 64u, (uint32_t)((z0 + 1) & 7) * 64u};
    uint32_t sl[8], raw[8];
    if (!__any(ox || oy || oz)) {
        // every active lane's corners lie in its brick b0: one address, the
        // eight loads at fixed offsets from it
        const uint32_t s0 = n.slot[0];
        const bool live = s0 < RKF_INTERIOR_SLOT;
        const uint16_t *pg = S.dist + (uint64_t)(live ? s0 : 0u) * 512u + (live ? lx[0] + ly[0] + lz[0] : 0u);
        const uint32_t off[8] = {0u, 1u, 8u, 9u, 64u, 65u, 72u, 73u};
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++) {
            sl[k] = s0;
            raw[k] = pg[off[k]];
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++) {
            const uint32_t kx = k & 1u, ky = (k >> 1) & 1u, kz = k >> 2;
            sl[k] = nb_pick(n, kx && ox, ky && oy, kz && oz);
            const bool live = sl[k] < RKF_INTERIOR_SLOT;
            const uint32_t loc = lx[kx] + ly[ky] + lz[kz];
            raw[k] = S.dist[(uint64_t)(live ? sl[k] : 0u) * 512u + (live ? loc : 0u)];
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) loaded(raw[k]);
    float c[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        const uint32_t sent = __float_as_uint(sl[k] == RKF_EMPTY_SLOT ? v_empty : v_interior);
        c[k] = __uint_as_float(sl[k] < RKF_INTERIOR_SLOT ? __float_as_uint(h2f(raw[k])) : sent);
    }
    const float c00 = mix(c[0], c[1], tx), c10 = mix(c[2], c[3], tx);
    const float c01 = mix(c[4], c[5], tx), c11 = mix(c[6], c[7], tx);
    const float c0 = mix(c00, c10, ty), c1 = mix(c01, c11, ty);
    return mix(c0, c1, tz) + 0.0f;
}

// In-grid test of sp_free: g and v0 of the point, false near the grid faces
// (those points take sp_sample's clamped path).
RKF_DEV bool sp_grid_pos(const SparseDev &S, f3 pos, float g[3], int32_t v0[3]) {
    const float vs = S.vs;
    const float be = vs * 8.0f;
    const f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    const f3 gp = sub(pos, mk(S.origin[0], S.origin[1], S.origin[2]));
    const float lo = vs * 0.01f;
    // one combined test (no early exits, so no nested branches)
    bool in = !(gp.x < lo) & !(gp.y < lo) & !(gp.z < lo) & !(gp.x > gs.x - lo) & !(gp.y > gs.y - lo) &
              !(gp.z > gs.z - lo);
    g[0] = gp.x * S.inv_vs - 0.5f;
    g[1] = gp.y * S.inv_vs - 0.5f;
    g[2] = gp.z * S.inv_vs - 0.5f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        v0[a] = (int32_t)floorf(g[a]);
        in = in & (v0[a] >= 0) & (v0[a] <= S.tot[a] - 2);
    }
    return in;
}

RKF_DEV uint64_t nb_key(const SparseDev &S, const int32_t v0[3]) {
    // brick coordinates of v0 packed 21 bits each (shifts only, no multiply)
    const uint64_t bx = (uint32_t)v0[0] >> 3, by = (uint32_t)v0[1] >> 3, bz = (uint32_t)v0[2] >> 3;
    return bx | (by << 21) | (bz << 42);
}

enum : int { kLocFree = 0, kLocCached = 1, kLocEdge = 2 };

// Low dword of a row word: the brick's L1 slot.
RKF_DEV const uint32_t *roww_slots(const SparseDev &S) { return reinterpret_cast<const uint32_t *>(S.roww); }

// Row of the cell containing corner brick k of a neighbourhood whose lower
// brick lies on the cell's +x / +y / +z face when fx / fy / fz: rows[0] is
// the cell's own, rows[j] the neighbour at offset (j & 1, j >> 1 & 1, j >> 2).
// Selects only, so the rows stay in registers.
RKF_DEV uint32_t nb_row_of(const uint32_t rows[8], uint32_t k, bool fx, bool fy, bool fz) {
    const bool ox = (k & 1u) && fx, oy = ((k >> 1) & 1u) && fy, oz = (k >> 2) && fz;
    const uint32_t a0 = ox ? rows[1] : rows[0], a1 = ox ? rows[3] : rows[2];
    const uint32_t a2 = ox ? rows[5] : rows[4], a3 = ox ? rows[7] : rows[6];
    const uint32_t b0 = oy ? a1 : a0, b1 = oy ? a3 : a2;
    return oz ? b1 : b0;
}

// Free-space test + exit parameter (see file header), with the brick
// neighbourhood cache. A point whose brick b0 is the cached one samples at
// once (a cached brick is never free). Otherwise the cell word decides
// cell-level skips (one load when the cell is not the ray's cached one,
// together with the L2 entries of the seven cells after it on +x, +y, +z),
// then the brick's row word brick-level skips; a brick that is not free
// refills the cache with the slots of its 2x2x2 neighbourhood, read from
// the row words of its own cell and of the neighbouring cells in one round.
// Near the grid faces the point takes the clamped sampling path (kLocEdge).
//
// The neighbour rows (L1 block, or EMPTY when the cell holds none) are kept
// per lane in LDS, tagged with their cell: they depend on the cell only, so
// any ray of the lane may reuse them. Kernels that call this run 64-lane
// workgroups.
RKF_DEV int sp_locate(const SparseDev &S, f3 pos, f3 sd, f3 inv, float t, float &t_exit, BrickNbhd &n, float g[3],
                      int32_t v0[3]) {
    __shared__ uint32_t lds_cell[64];
    __shared__ uint32_t lds_rows[7][64];
    if (!sp_grid_pos(S, pos, g, v0)) return kLocEdge;
    const uint64_t key = nb_key(S, v0);
    if (key == n.key) return kLocCached;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b[3] = {(uint32_t)v0[0] >> 3, (uint32_t)v0[1] >> 3, (uint32_t)v0[2] >> 3};
    const uint32_t q[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    const uint64_t cell = (uint64_t)q[0] + (uint64_t)q[1] * S.cells[0] + (uint64_t)q[2] * S.cells[0] * S.cells[1];
    const uint64_t cxy = (uint64_t)S.cells[0] * S.cells[1];
    float lo_b[3], hi_b[3];
    const uint32_t octant = (sd.x < 0.0f ? 1u : 0u) | (sd.y < 0.0f ? 2u : 0u) | (sd.z < 0.0f ? 4u : 0u);
    uint32_t e = 0u;
    uint32_t rows[8];
#pragma unroll
    for (uint32_t j = 0; j < 8u; j++) rows[j] = RKF_EMPTY_SLOT;
    bool have = false;
    if ((uint32_t)cell != n.cell) {
        const uint64_t cw = S.cellw[cell];
        uint32_t nl[8];
        bool ing[8];
#pragma unroll
        for (uint32_t j = 1; j < 8u; j++) {
            const uint32_t dx = j & 1u, dy = (j >> 1) & 1u, dz = j >> 2;
            ing[j] = q[0] + dx < S.cells[0] && q[1] + dy < S.cells[1] && q[2] + dz < S.cells[2];
            nl[j] = S.l2[ing[j] ? cell + dx + (uint64_t)dy * S.cells[0] + (uint64_t)dz * cxy : 0u];
        }
        uint32_t cw_lo = (uint32_t)cw, cw_hi = (uint32_t)(cw >> 32);
        loaded(cw_lo);
        loaded(cw_hi);
#pragma unroll
        for (uint32_t j = 1; j < 8u; j++) loaded(nl[j]);
        e = (cw_lo >> (4u * octant)) & 15u;
        if (e == 0) {
            n.cell = (uint32_t)cell;
            n.row = cw_hi;
            lds_cell[lane] = (uint32_t)cell;
#pragma unroll
            for (uint32_t j = 1; j < 8u; j++) {
                rows[j] = ing[j] ? nl[j] : RKF_EMPTY_SLOT;
                lds_rows[j - 1u][lane] = rows[j];
            }
            have = true;
        }
    }
    if (e > 0) {
        // the cube of cells q + [0, e) in the travel direction of each axis is free
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)q[a] - (float)(e - 1u)) * 64.0f;
            hi_b[a] = ((float)q[a] + (float)e) * 64.0f;
        }
    } else {
        // The brick's row word (extent and slot) and, side by side, the
        // slots of the rest of its 2x2x2 neighbourhood from the row words of
        // the cells they lie in, without branches so the loads go out
        // together. Unused loads read entry 0 (the arrays always hold one);
        // bricks past the grid end count EMPTY.
        const bool fx = (b[0] & 7u) == 7u, fy = (b[1] & 7u) == 7u, fz = (b[2] & 7u) == 7u;
        if (!have && (fx || fy || fz)) {
            if (lds_cell[lane] == (uint32_t)cell) {
#pragma unroll
                for (uint32_t j = 1; j < 8u; j++) rows[j] = lds_rows[j - 1u][lane];
            } else {
                uint32_t nl[8];
                bool ing[8];
#pragma unroll
                for (uint32_t j = 1; j < 8u; j++) {
                    const uint32_t dx = j & 1u, dy = (j >> 1) & 1u, dz = j >> 2;
                    ing[j] = q[0] + dx < S.cells[0] && q[1] + dy < S.cells[1] && q[2] + dz < S.cells[2];
                    nl[j] = S.l2[ing[j] ? cell + dx + (uint64_t)dy * S.cells[0] + (uint64_t)dz * cxy : 0u];
                }
                lds_cell[lane] = (uint32_t)cell;
#pragma unroll
                for (uint32_t j = 1; j < 8u; j++) {
                    loaded(nl[j]);
                    rows[j] = ing[j] ? nl[j] : RKF_EMPTY_SLOT;
                    lds_rows[j - 1u][lane] = rows[j];
                }
            }
        }
        rows[0] = n.row;
        const uint32_t lb = (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u;
        const uint64_t *pr = S.roww + (uint64_t)rows[0] * 512u + lb;
        const uint64_t rw = pr[0];
        uint32_t va[8];
        bool ok[8];
        if (!__any(fx || fy || fz)) {
            // no active lane's neighbourhood leaves its cell: all eight row
            // words at fixed offsets from the brick's own
            const uint32_t dk[8] = {0u, 1u, 8u, 9u, 64u, 65u, 72u, 73u};
#pragma unroll
            for (uint32_t k = 1; k < 8u; k++) {
                ok[k] = true;
                va[k] = *reinterpret_cast<const uint32_t *>(pr + dk[k]);
            }
        } else {
#pragma unroll
            for (uint32_t k = 1; k < 8u; k++) {
                const uint32_t bx = b[0] + (k & 1u), by = b[1] + ((k >> 1) & 1u), bz = b[2] + (k >> 2);
                const uint32_t rk = nb_row_of(rows, k, fx, fy, fz);
                ok[k] = bx < S.nb[0] && by < S.nb[1] && bz < S.nb[2] && rk != RKF_EMPTY_SLOT;
                const uint32_t lbk = (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u;
                va[k] = roww_slots(S)[ok[k] ? ((uint64_t)rk * 512u + lbk) * 2u : 0u];
            }
        }
        uint32_t rw_lo = (uint32_t)rw, rw_hi = (uint32_t)(rw >> 32);
        loaded(rw_lo);
        loaded(rw_hi);
#pragma unroll
        for (uint32_t k = 1; k < 8u; k++) loaded(va[k]);
        const uint32_t eb = (rw_hi >> (4u * octant)) & 15u;
        if (eb == 0) {
            n.key = key;
            n.slot[0] = rw_lo;
#pragma unroll
            for (uint32_t k = 1; k < 8u; k++) n.slot[k] = ok[k] ? va[k] : RKF_EMPTY_SLOT;
            return kLocCached;
        }
        // the cube of bricks b + [0, eb) in the travel direction of each axis is free
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)b[a] - (float)(eb - 1u)) * 8.0f;
            hi_b[a] = ((float)b[a] + (float)eb) * 8.0f;
        }
    }
    const float dv[3] = {sd.x, sd.y, sd.z};
    const float iv[3] = {inv.x, inv.y, inv.z};
    float dt = kMaxFloat;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float bound = dv[a] > 0.0f ? hi_b[a] : lo_b[a];
        const float da = (bound - g[a]) * iv[a];
        dt = fminf(dt, da);
    }
    t_exit = t + fmaxf(dt, 0.0f);
    return kLocFree;
}

// A stepping sample at any point (bisection): from the cache when the point's
// brick is the cached one, else sp_sample.
RKF_DEV float sp_sample_nb(SpCtx &c, const BrickNbhd &n, f3 p) {
    float g[3];
    int32_t v0[3];
    if (sp_grid_pos(*c.S, p, g, v0) && nb_key(*c.S, v0) == n.key) {
        c.evals++;
        return nb_sample(*c.S, n, g, v0);
    }
    return sp_sample(c, p, false, true);
}

// ---- Sub-wave groups: 8 lanes per ray ----
// Once at most kGroupRays rays of a wavefront are still marching, each is
// handed to a group of 8 lanes (lanes 8g..8g+7 take the g-th unfinished ray
// in lane order). Every lane of a group runs the ray's step logic on the same
// values; the step's loads are split across the group (lane k: the L2 entry
// of neighbour cell k, the slot of corner brick k, corner voxel k) and the
// trilinear mix is reduced across it by lane exchanges, with the operations
// and operand order of nb_sample, so a ray's results are those of its
// one-lane march. A lone ray then issues a step's 8 loads and their address
// arithmetic as one instruction each instead of eight, which shortens the
// wavefront's remaining steps (its few long rays set the launch's end).
// Each lane of a group loads one slot of the 2x2x2 neighbourhood (lane k:
// corner brick b0 + (k & 1, k >> 1 & 1, k >> 2)) and the group exchanges
// them once per refill, so every lane holds the whole cache and a sample
// picks its corner's slot by selects; lane k >= 1 keeps the L1 block of
// neighbour cell (k & 1, k >> 1 & 1, k >> 2) of nb.cell in nb.nrow.
constexpr uint32_t kGroupRays = 8u;

// sp_locate for a group (lane k of the group whose first lane is `base`):
// same decisions, same cache contents (spread as above).
RKF_DEV int sp_locate_g(const SparseDev &S, f3 pos, f3 sd, f3 inv, float t, float &t_exit, BrickNbhd &n,
                        float g[3], int32_t v0[3], uint32_t k, uint32_t base) {
    if (!sp_grid_pos(S, pos, g, v0)) return kLocEdge;
    const uint64_t key = nb_key(S, v0);
    if (key == n.key) return kLocCached;
    const uint32_t kx = k & 1u, ky = (k >> 1) & 1u, kz = k >> 2;
    const uint32_t b[3] = {(uint32_t)v0[0] >> 3, (uint32_t)v0[1] >> 3, (uint32_t)v0[2] >> 3};
    const uint32_t q[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    const uint64_t cell = (uint64_t)q[0] + (uint64_t)q[1] * S.cells[0] + (uint64_t)q[2] * S.cells[0] * S.cells[1];
    const uint64_t cxy = (uint64_t)S.cells[0] * S.cells[1];
    float lo_b[3], hi_b[3];
    const uint32_t octant = (sd.x < 0.0f ? 1u : 0u) | (sd.y < 0.0f ? 2u : 0u) | (sd.z < 0.0f ? 4u : 0u);
    uint32_t e = 0u;
    if ((uint32_t)cell != n.cell) {
        // every lane: the cell word (one address); lane k >= 1: the L2 entry of neighbour cell k
        const bool ing = k != 0u && q[0] + kx < S.cells[0] && q[1] + ky < S.cells[1] && q[2] + kz < S.cells[2];
        const uint64_t cw = S.cellw[cell];
        uint32_t nl = S.l2[ing ? cell + kx + (uint64_t)ky * S.cells[0] + (uint64_t)kz * cxy : 0u];
        uint32_t cw_lo = (uint32_t)cw, cw_hi = (uint32_t)(cw >> 32);
        loaded(cw_lo);
        loaded(cw_hi);
        loaded(nl);
        e = (cw_lo >> (4u * octant)) & 15u;
        if (e == 0) {
            n.cell = (uint32_t)cell;
            n.row = cw_hi;
            n.nrow = ing ? nl : RKF_EMPTY_SLOT;
        }
    }
    if (e > 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)q[a] - (float)(e - 1u)) * 64.0f;
            hi_b[a] = ((float)q[a] + (float)e) * 64.0f;
        }
    } else {
        // lane k: the slot of corner brick k from the row word in the row of
        // the cell it lies in (the neighbour cell's, read from lane j of the
        // group, when it crosses a face); every lane: the brick's own row word
        const bool fx = (b[0] & 7u) == 7u, fy = (b[1] & 7u) == 7u, fz = (b[2] & 7u) == 7u;
        const uint32_t j = ((kx && fx) ? 1u : 0u) | ((ky && fy) ? 2u : 0u) | ((kz && fz) ? 4u : 0u);
        const uint32_t nrow = lane_read(n.nrow, base + j);
        const uint32_t rk = j == 0u ? n.row : nrow;
        const uint32_t bx = b[0] + kx, by = b[1] + ky, bz = b[2] + kz;
        const bool ok = bx < S.nb[0] && by < S.nb[1] && bz < S.nb[2] && rk != RKF_EMPTY_SLOT;
        const uint32_t lb = (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u;
        const uint32_t lbk = (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u;
        const uint64_t rw = S.roww[(uint64_t)n.row * 512u + lb];
        uint32_t va = roww_slots(S)[ok ? ((uint64_t)rk * 512u + lbk) * 2u : 0u];
        uint32_t rw_lo = (uint32_t)rw, rw_hi = (uint32_t)(rw >> 32);
        loaded(rw_lo);
        loaded(rw_hi);
        loaded(va);
        const uint32_t eb = (rw_hi >> (4u * octant)) & 15u;
        if (eb == 0) {
            n.key = key;
            const uint32_t mine = ok ? va : RKF_EMPTY_SLOT;
#pragma unroll
            for (uint32_t m = 0; m < 8u; m++) n.slot[m] = lane_read(mine, base + m);
            return kLocCached;
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((float)b[a] - (float)(eb - 1u)) * 8.0f;
            hi_b[a] = ((float)b[a] + (float)eb) * 8.0f;
        }
    }
    const float dv[3] = {sd.x, sd.y, sd.z};
    const float iv[3] = {inv.x, inv.y, inv.z};
    float dt = kMaxFloat;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float bound = dv[a] > 0.0f ? hi_b[a] : lo_b[a];
        const float da = (bound - g[a]) * iv[a];
        dt = fminf(dt, da);
    }
    t_exit = t + fmaxf(dt, 0.0f);
    return kLocFree;
}

// nb_sample for a group: lane k loads corner k (k = x + 2y + 4z), the
// group reduces x, then y, then z with nb_sample's mixes.
RKF_DEV float nb_sample_g(const SparseDev &S, const BrickNbhd &n, const float g[3], const int32_t v0[3], uint32_t k,
                          uint32_t base) {
    const float vs = S.vs;
    const float v_empty = vs * 8.0f, v_interior = -(vs * 2.0f);
    const int32_t x0 = v0[0], y0 = v0[1], z0 = v0[2];
    const float tx = g[0] - (float)x0, ty = g[1] - (float)y0, tz = g[2] - (float)z0;
    const uint32_t kx = k & 1u, ky = (k >> 1) & 1u, kz = k >> 2;
    const bool ox = (x0 & 7) == 7, oy = (y0 & 7) == 7, oz = (z0 & 7) == 7;
    const uint32_t sl = nb_pick(n, kx && ox, ky && oy, kz && oz);
    const bool live = sl < RKF_INTERIOR_SLOT;
    const uint32_t loc = (uint32_t)((x0 + (int32_t)kx) & 7) + (uint32_t)((y0 + (int32_t)ky) & 7) * 8u +
                         (uint32_t)((z0 + (int32_t)kz) & 7) * 64u;
    uint32_t raw = S.dist[(uint64_t)(live ? sl : 0u) * 512u + (live ? loc : 0u)];
    loaded(raw);
    const uint32_t sent = __float_as_uint(sl == RKF_EMPTY_SLOT ? v_empty : v_interior);
    const float c = __uint_as_float(live ? __float_as_uint(h2f(raw)) : sent);
    const float p1 = dpp_xor1(c);
    const float c1 = kx ? mix(p1, c, tx) : mix(c, p1, tx);
    const float p2 = dpp_xor2(c1);
    const float c2 = ky ? mix(p2, c1, ty) : mix(c1, p2, ty);
    const float p3 = dpp_half_mirror(c2);
    return (kz ? mix(p3, c2, tz) : mix(c2, p3, tz)) + 0.0f;
}

struct SpRes {
    float t, mat_d;
    f3 mat_p;
    bool hit;
    uint32_t steps;
};

// Resumable primary-ray state: one iteration of ray_march_tiled's loop
// (:905-982) per prim_step call, so the plain kernel (loop to completion) and
// the tuned kernel (interleaved with lane refills) run identical arithmetic.
struct PrimRay {
    f3 o, sd;  // origin, safe direction (shadow rays: light direction)
    f3 inv;    // vs / sd per axis: free-space exits in voxel units -> t
    float t, prev_t, prev_dist;  // shadow rays keep their visibility in prev_dist
    float tmax;                  // max_distance, or less with a tile interval
    uint32_t near, steps, evals;
    BrickNbhd nb;                // slots around the last sampled brick
};
enum : int { kRayContinue = 0, kRayHit = 1, kRayMiss = 2 };

RKF_DEV void prim_init(PrimRay &r, f3 o, f3 dir, float vs, float max_distance) {
    r.o = o;
    r.sd = safe_dir(dir);
    r.inv = mk(vs / r.sd.x, vs / r.sd.y, vs / r.sd.z);
    r.t = 0.0f;
    r.prev_t = 0.0f;
    r.prev_dist = kMaxFloat;
    r.tmax = max_distance;
    r.near = 0;
    r.steps = 0;
    r.evals = 0;
    nb_reset(r.nb);
}

// Tile interval (t_lo, t_hi): outside it every sample along the ray reads
// the free value, so the ray skips to t_lo (one iteration, as a free-space
// skip) and stops as a miss past t_hi.
RKF_DEV void prim_cull(PrimRay &r, uint2 iv, float vs) {
    const float lo = __uint_as_float(iv.x), hi = __uint_as_float(iv.y);
    r.tmax = fminf(r.tmax, hi);
    if (lo > r.t) {
        r.steps = 1;
        r.prev_t = lo;
        r.prev_dist = vs * 8.0f;
        r.t = lo + vs * 0.00390625f;
    }
}

// Step codes: kModeThroughput (sp_free + sp_sample with the prefetched
// slot: fewer memory instructions, for launches with enough rays in flight
// to hide latency), kModeLowLat (the brick-neighbourhood cache, sp_locate +
// nb_sample: fewer dependent loads per step, for launches bounded by their
// slowest rays), kModeGroup (kModeLowLat spread over an 8-lane group). All
// give the same values.
enum : int { kModeThroughput = 0, kModeLowLat = 1, kModeGroup = 2 };

// One step's field query: false for a free-space skip (t_exit set), else
// the stepping sample d.
template <int kMode>
RKF_DEV bool step_field(const SparseDev &S, PrimRay &r, f3 pos, float &t_exit, SpCtx &c, float &d) {
    if constexpr (kMode != kModeThroughput) {
        float g[3];
        int32_t v0[3];
        const uint32_t k = threadIdx.x & 7u, base = (threadIdx.x & 63u) & ~7u;
        int loc;
        if constexpr (kMode == kModeGroup) loc = sp_locate_g(S, pos, r.sd, r.inv, r.t, t_exit, r.nb, g, v0, k, base);
        else loc = sp_locate(S, pos, r.sd, r.inv, r.t, t_exit, r.nb, g, v0);
        if (loc == kLocFree) return false;
        if (loc == kLocCached) {
            c.evals++;
            if constexpr (kMode == kModeGroup) d = nb_sample_g(S, r.nb, g, v0, k, base);
            else d = nb_sample(S, r.nb, g, v0);
        } else {
            d = sp_sample(c, pos, false, true);
        }
        return true;
    } else {
        uint32_t slot;
        if (sp_free(S, pos, r.sd, r.inv, r.t, t_exit, slot)) return false;
        d = sp_sample(c, pos, false, true, slot);
        return true;
    }
}

template <int kMode>
RKF_DEV int prim_step(const SparseDev &S, PrimRay &r, const rkf_scene_uniforms &su, uint32_t eps_mode, SpRes &out) {
    if (r.steps >= kHardMax || r.t > r.tmax || r.near >= su.max_steps) return kRayMiss;
    const float vs = S.vs;
    r.steps++;
    f3 pos = add(r.o, muls(r.sd, r.t));
    float tx;
    SpCtx c{&S, r.evals};
    float d;
    if (!step_field<kMode>(S, r, pos, tx, c, d)) {
        r.prev_t = tx;
        r.prev_dist = vs * 8.0f;
        r.t = tx + vs * 0.00390625f;
        return kRayContinue;
    }
    if (d < 0.0f && r.prev_dist > hit_eps(eps_mode, su.hit_threshold, r.prev_t)) {
        float lo = r.prev_t, hi = r.t, mt = r.t, md = d;
        // low latency: the bricks within 3 voxels of the segment's midpoint,
        // looked up once for the 8 bisection samples
        BrickCache bc;
        if constexpr (kMode != kModeThroughput) bc_init(S, add(r.o, muls(r.sd, (lo + hi) * 0.5f)), 3, bc);
        for (uint32_t b = 0; b < kBisect; b++) {
            float mid = (lo + hi) * 0.5f;
            const f3 mp = add(r.o, muls(r.sd, mid));
            float mdd;
            if constexpr (kMode != kModeThroughput) {
                c.evals++;
                mdd = sp_sample_bc(S, bc, mp, false);
            } else {
                mdd = sp_sample(c, mp, false, true);
            }
            if (mdd < 0.0f) {
                hi = mid;
            } else {
                lo = mid;
                mt = mid;
                md = mdd;
            }
        }
        r.evals = c.evals;
        out.t = (lo + hi) * 0.5f;
        out.hit = true;
        out.mat_p = add(r.o, muls(r.sd, mt));
        out.mat_d = md;
        return kRayHit;
    }
    r.evals = c.evals;
    if (d < hit_eps(eps_mode, su.hit_threshold, r.t)) {
        out.t = r.t;
        out.hit = true;
        out.mat_p = pos;
        out.mat_d = d;
        return kRayHit;
    }
    if (d < kNear) r.near++;
    r.prev_t = r.t;
    r.prev_dist = d;
    r.t += fmaxf(d, kMinStep);
    return kRayContinue;
}

// Resumable soft shadow (shade_common_shading.wgsl:9-33) in the same state
// record: sd = light direction, prev_dist = running visibility.
using ShadowRay = PrimRay;

RKF_DEV void shadow_init(ShadowRay &r, f3 o, f3 L, float vs) {
    r.o = o;
    r.sd = L;
    r.inv = mk(vs / L.x, vs / L.y, vs / L.z);
    r.t = 0.005f;
    r.prev_t = 0.0f;
    r.prev_dist = 1.0f;
    r.near = 0;
    r.steps = 0;
    r.evals = 0;
    nb_reset(r.nb);
}

// Returns true when finished; r.prev_dist then holds the visibility.
template <int kMode>
RKF_DEV bool shadow_step(const SparseDev &S, ShadowRay &r, float maxd, float k) {
    if (r.steps >= 256u || r.near >= 32u || r.t > maxd) {
        r.prev_dist = clampf(r.prev_dist, 0.0f, 1.0f);
        return true;
    }
    r.steps++;
    f3 p = add(r.o, muls(r.sd, r.t));
    float tx;
    SpCtx c{&S, r.evals};
    float d;
    if (!step_field<kMode>(S, r, p, tx, c, d)) {
        r.t = tx + S.vs * 0.00390625f;
        return false;
    }
    r.evals = c.evals;
    if (d < 0.005f) {
        r.prev_dist = 0.0f;
        return true;
    }
    if (d < 0.02f) {
        r.prev_dist = fminf(r.prev_dist, k * d / r.t);
        r.near++;
    }
    r.t += fmaxf(d, 0.005f);
    return false;
}

RKF_DEV float shadow_maxd(const SparseMarchParams &P) { return P.ext.shadow_max_dist > 0.0f ? P.ext.shadow_max_dist : 25.0f; }
RKF_DEV float shadow_k(const SparseMarchParams &P) { return P.ext.shadow_k > 0.0f ? P.ext.shadow_k : 16.0f; }

// Writes the G-buffer (and debug record without shadow counters) for a
// finished primary ray; returns true and the shadow-ray origin when a shadow
// ray is due.
RKF_DEV bool prim_finish(const SparseMarchParams &P, uint32_t px, uint32_t py, const PrimRay &r, int status,
                         const SpRes &res, f3 &shadow_o) {
    const uint32_t W = P.g.width, H = P.g.height;
    uint64_t p = (uint64_t)py * W + px;
    if (status != kRayHit) {
        store_miss(P.g, p);
        store_debug(P.g, p, mk(0.0f, 0.0f, 0.0f), RKF_EMPTY_SLOT, 0xFFFFFFFFu, 0xFFFFFFFFu, r.steps, r.evals);
        if (P.g.shadow) P.g.shadow[p] = 0.0f;
        return false;
    }
    const SparseDev &S = P.S;
    f3 o, dir;
    camera_ray(P.cam, W, H, px, py, o, dir);
    f3 hp = add(o, muls(dir, res.t));
    float e = S.vs * 2.0f;
    BrickCache bc;
    bc_init(S, hp, 3, bc);
    uint32_t mat = 0;
    if (res.mat_d < S.vs * 16.0f) mat = sp_material(S, res.mat_p, bc);
    float gx = sp_sample_bc(S, bc, add(hp, mk(e, 0.0f, 0.0f)), true) - sp_sample_bc(S, bc, sub(hp, mk(e, 0.0f, 0.0f)), true);
    float gy = sp_sample_bc(S, bc, add(hp, mk(0.0f, e, 0.0f)), true) - sp_sample_bc(S, bc, sub(hp, mk(0.0f, e, 0.0f)), true);
    float gz = sp_sample_bc(S, bc, add(hp, mk(0.0f, 0.0f, e)), true) - sp_sample_bc(S, bc, sub(hp, mk(0.0f, 0.0f, e)), true);
    f3 gv = mk(gx, gy, gz);
    float len = length(gv);
    f3 n = len < 1e-10f ? mk(0.0f, 1.0f, 0.0f) : mk(gv.x / len, gv.y / len, gv.z / len);
    uint32_t slot, vox, cell;
    uint64_t vi = sp_nearest(S, hp, slot, vox, cell, bc);
    uint32_t sec = 0, blend = 0;
    if (S.aux && vi != ~0ull) {
        uint32_t a = S.aux[vi];
        sec = a & 0x3Fu;
        blend = (a >> 8) & 0xFFu;
    }
    float sdf_at = sp_sample_bc(S, bc, hp, false);
    uint32_t packed = (mat & 0x3Fu) | (sec << 6) | (blend << 12) | (P.object_id << 24);
    float mx, my;
    motion_vec(P.cam, W, H, px, py, hp, mx, my);
    store_hit(P.g, p, hp, res.t, n, packed, mx, my, sdf_at);
    store_debug(P.g, p, n, slot, vox, cell, r.steps, r.evals);
    if (!P.g.shadow) return false;
    if (!P.ext.shadow_enable) {
        P.g.shadow[p] = 0.0f;
        return false;
    }
    // The shading pass reads the normal back from the rgba16f target.
    f3 nq = mk(h2f(f2h(n.x)), h2f(f2h(n.y)), h2f(f2h(n.z)));
    f3 L = mk(P.ext.light_dir[0], P.ext.light_dir[1], P.ext.light_dir[2]);
    const float bias = 0.08f;
    shadow_o = add(add(hp, muls(nq, bias)), muls(L, bias * 0.5f));
    return true;
}

RKF_DEV void shadow_store(const SparseMarchParams &P, uint64_t p, const ShadowRay &s) {
    P.g.shadow[p] = s.prev_dist;
    if (P.g.debug) {
        P.g.debug[p].shadow_steps = s.steps;
        P.g.debug[p].shadow_evals = s.evals;
    }
}

// Group of lane `lane` when the unfinished rays are the set bits of `pend`
// (wave-uniform): the lane holding the group's ray (the g-th set bit).
RKF_DEV uint32_t group_source(uint64_t pend, uint32_t grp) {
    uint32_t src = 0u;
    uint64_t m = pend;
#pragma unroll
    for (uint32_t i = 0; i < kGroupRays; i++) {
        const uint32_t b = (uint32_t)__builtin_ctzll(m | 0x8000000000000000ull);
        src = grp == i ? b : src;
        m &= m - 1ull;
    }
    return src;
}

RKF_DEV float lane_readf(float v, uint32_t src) { return __uint_as_float(lane_read(__float_as_uint(v), src)); }
RKF_DEV f3 lane_read3(f3 v, uint32_t src) { return mk(lane_readf(v.x, src), lane_readf(v.y, src), lane_readf(v.z, src)); }

// The ray state a group needs from the ray's own lane (the neighbourhood
// cache starts empty: it never changes a value).
RKF_DEV void group_gather(const PrimRay &r, uint32_t src, PrimRay &q) {
    q.o = lane_read3(r.o, src);
    q.sd = lane_read3(r.sd, src);
    q.inv = lane_read3(r.inv, src);
    q.t = lane_readf(r.t, src);
    q.prev_t = lane_readf(r.prev_t, src);
    q.prev_dist = lane_readf(r.prev_dist, src);
    q.tmax = lane_readf(r.tmax, src);
    q.near = lane_read(r.near, src);
    q.steps = lane_read(r.steps, src);
    q.evals = lane_read(r.evals, src);
    nb_reset(q.nb);
}

// Primary rays still marching (st == kRayContinue) when at most kGroupRays
// of the wave are: finished in 8-lane groups, results back in their lanes.
// All 64 lanes of the wavefront must be active.
RKF_DEV void prim_groups(const SparseMarchParams &P, PrimRay &r, int &st, SpRes &res) {
    const uint64_t pend = __ballot(st == kRayContinue);
    if (pend == 0ull) return;
    const uint32_t lane = threadIdx.x & 63u, grp = lane >> 3;
    PrimRay q;
    group_gather(r, group_source(pend, grp), q);
    SpRes qr;
    qr.t = qr.mat_d = 0.0f;
    qr.mat_p = mk(0.0f, 0.0f, 0.0f);
    qr.hit = false;
    qr.steps = 0u;
    int qs = kRayContinue;
    if (grp < (uint32_t)__popcll(pend)) {
        do {
            qs = prim_step<kModeGroup>(P.S, q, P.su, P.ext.hit_eps_mode, qr);
        } while (qs == kRayContinue);
    }
    const uint32_t lead = 8u * (uint32_t)__popcll(pend & ((1ull << lane) - 1ull));
    const int s2 = (int)lane_read((uint32_t)qs, lead);
    const uint32_t steps = lane_read(q.steps, lead), evals = lane_read(q.evals, lead);
    SpRes r2;
    r2.t = lane_readf(qr.t, lead);
    r2.mat_d = lane_readf(qr.mat_d, lead);
    r2.mat_p = lane_read3(qr.mat_p, lead);
    r2.hit = lane_read(qr.hit ? 1u : 0u, lead) != 0u;
    r2.steps = lane_read(qr.steps, lead);
    if ((pend >> lane) & 1ull) {
        st = s2;
        r.steps = steps;
        r.evals = evals;
        res = r2;
    }
}

// Shadow rays still marching (pending) when at most kGroupRays of the wave
// are: finished in 8-lane groups, visibility and counters back in their
// lanes. All 64 lanes of the wavefront must be active.
RKF_DEV void shadow_groups(const SparseDev &S, ShadowRay &r, bool pending, float maxd, float k) {
    const uint64_t pend = __ballot(pending);
    if (pend == 0ull) return;
    const uint32_t lane = threadIdx.x & 63u, grp = lane >> 3;
    ShadowRay q;
    group_gather(r, group_source(pend, grp), q);
    if (grp < (uint32_t)__popcll(pend)) {
        while (!shadow_step<kModeGroup>(S, q, maxd, k)) {
        }
    }
    const uint32_t lead = 8u * (uint32_t)__popcll(pend & ((1ull << lane) - 1ull));
    const float vis = lane_readf(q.prev_dist, lead);
    const uint32_t steps = lane_read(q.steps, lead), evals = lane_read(q.evals, lead);
    if (pending) {
        r.prev_dist = vis;
        r.steps = steps;
        r.evals = evals;
    }
}

// One pixel of a view: its primary ray, the hit's shading and G-buffer
// record, then its shadow ray. `live` false: a lane with no pixel (past the
// view's edge), which only takes part in the wave's group hand-offs. With
// the low-latency step code, the last kGroupRays primary rays and the last
// kGroupRays shadow rays of the wavefront finish in 8-lane groups (all 64
// lanes of the wavefront must call this).
template <bool kLowLat>
RKF_DEV void view_pixel(const SparseMarchParams &P, uint32_t px, uint32_t py, bool live = true) {
    uint2 iv = make_uint2(0u, 0u);
    bool run = live;
    if (run && P.tiles) {
        iv = P.tiles[(uint64_t)(py >> 3) * P.tiles_x + (px >> 3)];
        if (iv.x == 0x7F800000u) {
            // a tile no occupied cell projects to: prim_cull sets t = +inf
            // and the first prim_step misses, so write that miss directly
            const uint64_t p = (uint64_t)py * P.g.width + px;
            store_miss(P.g, p);
            store_debug(P.g, p, mk(0.0f, 0.0f, 0.0f), RKF_EMPTY_SLOT, 0xFFFFFFFFu, 0xFFFFFFFFu, 1u, 0u);
            if (P.g.shadow) P.g.shadow[p] = 0.0f;
            run = false;
        }
    }
    if constexpr (!kLowLat) {
        if (!run) return;
        f3 o, dir;
        camera_ray(P.cam, P.g.width, P.g.height, px, py, o, dir);
        PrimRay r;
        prim_init(r, o, dir, P.S.vs, P.su.max_distance);
        if (P.tiles) prim_cull(r, iv, P.S.vs);
        SpRes res;
        int st;
        do {
            st = prim_step<kModeThroughput>(P.S, r, P.su, P.ext.hit_eps_mode, res);
        } while (st == kRayContinue);
        f3 so;
        if (prim_finish(P, px, py, r, st, res, so)) {
            ShadowRay sr;
            shadow_init(sr, so, mk(P.ext.light_dir[0], P.ext.light_dir[1], P.ext.light_dir[2]), P.S.vs);
            float maxd = shadow_maxd(P), k = shadow_k(P);
            while (!shadow_step<kModeThroughput>(P.S, sr, maxd, k)) {
            }
            shadow_store(P, (uint64_t)py * P.g.width + px, sr);
        }
    } else {
        PrimRay r;
        SpRes res;
        int st = kRayMiss;
        if (run) {
            f3 o, dir;
            camera_ray(P.cam, P.g.width, P.g.height, px, py, o, dir);
            prim_init(r, o, dir, P.S.vs, P.su.max_distance);
            if (P.tiles) prim_cull(r, iv, P.S.vs);
            for (;;) {
                st = prim_step<kModeLowLat>(P.S, r, P.su, P.ext.hit_eps_mode, res);
                const bool more = st == kRayContinue;
                const uint64_t m = __ballot(more);
                if (!more || (uint32_t)__popcll(m) <= kGroupRays) break;
            }
        }
        prim_groups(P, r, st, res);
        f3 so;
        const bool sh = run && prim_finish(P, px, py, r, st, res, so);
        const float maxd = shadow_maxd(P), k = shadow_k(P);
        ShadowRay sr;
        bool done = true;
        if (sh) {
            shadow_init(sr, so, mk(P.ext.light_dir[0], P.ext.light_dir[1], P.ext.light_dir[2]), P.S.vs);
            for (;;) {
                done = shadow_step<kModeLowLat>(P.S, sr, maxd, k);
                const uint64_t m = __ballot(!done);
                if (done || (uint32_t)__popcll(m) <= kGroupRays) break;
            }
        }
        shadow_groups(P.S, sr, !done, maxd, k);
        if (sh) shadow_store(P, (uint64_t)py * P.g.width + px, sr);
    }
}

// Point query (sdf_collision.rs:38-56): distance, nearest material, inside
// flag, central-difference gradient at eps 0.01 with normalize_or_zero
// (v * (1/len) when that reciprocal is finite and > 0).
// The 7 samples and the material share one brick-cache lookup around p (a
// corner outside it falls back to the index), each sample one round of loads.
RKF_DEV void point_query(const SparseDev &S, f3 p, rkfm_point_result &out) {
    BrickCache bc;
    bc_init(S, p, 3, bc);
    float d = sp_sample_bc(S, bc, p, false);
    const float e = 0.01f;
    float dx = sp_sample_bc(S, bc, add(p, mk(e, 0.0f, 0.0f)), false) - sp_sample_bc(S, bc, sub(p, mk(e, 0.0f, 0.0f)), false);
    float dy = sp_sample_bc(S, bc, add(p, mk(0.0f, e, 0.0f)), false) - sp_sample_bc(S, bc, sub(p, mk(0.0f, e, 0.0f)), false);
    float dz = sp_sample_bc(S, bc, add(p, mk(0.0f, 0.0f, e)), false) - sp_sample_bc(S, bc, sub(p, mk(0.0f, 0.0f, e)), false);
    float len = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    float rcp = 1.0f / len;
    rkfm_point_result r;
    r.distance = d;
    r.material = sp_material(S, p, bc);
    r.inside = d < 0.0f ? 1u : 0u;
    if (isfinite(rcp) && rcp > 0.0f) {
        r.gradient[0] = dx * rcp;
        r.gradient[1] = dy * rcp;
        r.gradient[2] = dz * rcp;
    } else {
        r.gradient[0] = r.gradient[1] = r.gradient[2] = 0.0f;
    }
    out = r;
}

// Sphere casts (declared with rkfm_sparse_sphere_cast): L(p), a lower bound
// on the distance from p to any stored brick, from the cell word's octant
// extents (free cells) or the brick's row word (free bricks). A free brick b
// has every brick of {b..b+1}^3 EMPTY, so the free box [b - E + 1, b + E - 1]
// (E = the minimum of the 8 octant extents) leaves [b - E + 1, b + E] EMPTY;
// likewise for cells.
RKF_DEV float sp_free_bound(const SparseDev &S, f3 p) {
    const float vs = S.vs, be = vs * 8.0f;
    const f3 gs = mk((float)S.nb[0] * be, (float)S.nb[1] * be, (float)S.nb[2] * be);
    const f3 gp = sub(p, mk(S.origin[0], S.origin[1], S.origin[2]));
    if (gp.x < 0.0f || gp.y < 0.0f || gp.z < 0.0f || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z) {
        const float dx = fmaxf(fmaxf(-gp.x, gp.x - gs.x), 0.0f), dy = fmaxf(fmaxf(-gp.y, gp.y - gs.y), 0.0f),
                    dz = fmaxf(fmaxf(-gp.z, gp.z - gs.z), 0.0f);
        return __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    }
    const float x[3] = {gp.x * S.inv_vs, gp.y * S.inv_vs, gp.z * S.inv_vs};
    uint32_t b[3];
#pragma unroll
    for (int a = 0; a < 3; a++) b[a] = (uint32_t)clampi((int32_t)floorf(x[a]), 0, S.tot[a] - 1) >> 3;
    const uint64_t cell = (uint64_t)(b[0] >> 3) + (uint64_t)(b[1] >> 3) * S.cells[0] +
                          (uint64_t)(b[2] >> 3) * S.cells[0] * S.cells[1];
    const uint64_t cw = S.cellw[cell];
    uint32_t e = 15u;
#pragma unroll
    for (uint32_t o = 0; o < 8u; o++) e = min(e, ((uint32_t)cw >> (4u * o)) & 15u);
    float unit = 64.0f;
    uint32_t base[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    if (e == 0u) {
        const uint32_t row = (uint32_t)(cw >> 32);
        if (row == kFreeCell) return 0.0f;
        const uint64_t rw = S.roww[(uint64_t)row * 512u + (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u];
        e = 15u;
#pragma unroll
        for (uint32_t o = 0; o < 8u; o++) e = min(e, (uint32_t)(rw >> (32u + 4u * o)) & 15u);
        if (e == 0u) return 0.0f;
        unit = 8.0f;
        base[0] = b[0];
        base[1] = b[1];
        base[2] = b[2];
    }
    float m = kMaxFloat;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float lo = ((float)base[a] - (float)(e - 1u)) * unit, hi = ((float)base[a] + (float)e) * unit;
        m = fminf(m, fminf(x[a] - lo, hi - x[a]));
    }
    return fmaxf(m, 0.0f) * vs;
}

RKF_DEV void sphere_cast(const SparseDev &S, const rkfm_sphere_cast &q, rkfm_cast_result &out) {
    const f3 o = mk(q.origin[0], q.origin[1], q.origin[2]), dir = mk(q.dir[0], q.dir[1], q.dir[2]);
    const float r = q.radius, vs = S.vs;
    SpCtx c{&S, 0u};
    rkfm_cast_result res;
    res.t = q.max_t;
    res.hit = 0u;
    res.steps = 0u;
    res.normal[0] = res.normal[1] = res.normal[2] = 0.0f;
    res.distance = 0.0f;
    res._pad = 0u;
    float t = 0.0f;
    for (uint32_t i = 0; i < kHardMax; i++) {
        if (t > q.max_t) break;
        res.steps++;
        const f3 p = add(o, muls(dir, t));
        const float L = sp_free_bound(S, p);
        const float D = L >= vs * 16.0f ? L : fmaxf(sp_sample(c, p, false, true), L);
        res.distance = D;
        if (D - r < 1e-3f) {
            const float e = vs * 2.0f;
            const float gx = sp_sample(c, add(p, mk(e, 0.0f, 0.0f)), true, false) -
                             sp_sample(c, sub(p, mk(e, 0.0f, 0.0f)), true, false);
            const float gy = sp_sample(c, add(p, mk(0.0f, e, 0.0f)), true, false) -
                             sp_sample(c, sub(p, mk(0.0f, e, 0.0f)), true, false);
            const float gz = sp_sample(c, add(p, mk(0.0f, 0.0f, e)), true, false) -
                             sp_sample(c, sub(p, mk(0.0f, 0.0f, e)), true, false);
            const float len = length(mk(gx, gy, gz));
            const f3 n = len < 1e-10f ? mk(0.0f, 1.0f, 0.0f) : mk(gx / len, gy / len, gz / len);
            res.t = t;
            res.hit = 1u;
            res.normal[0] = n.x;
            res.normal[1] = n.y;
            res.normal[2] = n.z;
            break;
        }
        t += fmaxf(D - r, kMinStep);
    }
    out = res;
}

// CollisionShape::sample_points (sdf_collision.rs:79-148): sample k of a
// body's shape in body-local space; false when the shape has no sample k.
RKF_DEV bool body_sample(const rkfm_body &b, uint32_t k, f3 &p) {
    if (b.shape == RKFM_BODY_SPHERE) {
        if (k >= 14u) return false;
        const float r = b.dims[0];
        if (k < 6u) {
            const float v = (k & 1u) ? -r : r;
            p = k < 2u ? mk(v, 0.0f, 0.0f) : (k < 4u ? mk(0.0f, v, 0.0f) : mk(0.0f, 0.0f, v));
            return true;
        }
        const float d3 = r / __builtin_sqrtf(3.0f);
        const uint32_t j = k - 6u;  // signs in x, y, z order: +++ ++- +-+ +-- -++ -+- --+ ---
        p = mk((j & 4u) ? -d3 : d3, (j & 2u) ? -d3 : d3, (j & 1u) ? -d3 : d3);
        return true;
    }
    if (b.shape == RKFM_BODY_BOX) {
        if (k >= 14u) return false;
        const float hx = b.dims[0], hy = b.dims[1], hz = b.dims[2];
        if (k < 8u) {  // corners: sx outer, sy, sz inner, -1 before +1
            p = mk((k & 4u) ? hx : -hx, (k & 2u) ? hy : -hy, (k & 1u) ? hz : -hz);
            return true;
        }
        const uint32_t j = k - 8u;  // face centres +x -x +y -y +z -z
        const float v = j < 2u ? hx : (j < 4u ? hy : hz);
        const float sv = (j & 1u) ? -v : v;
        p = j < 2u ? mk(sv, 0.0f, 0.0f) : (j < 4u ? mk(0.0f, sv, 0.0f) : mk(0.0f, 0.0f, sv));
        return true;
    }
    if (b.shape == RKFM_BODY_CAPSULE) {
        if (k >= 12u) return false;
        const float hh = b.dims[0], r = b.dims[1];
        if (k < 10u) {  // each hemisphere: pole, +x, -x, +z, -z
            const float y = k < 5u ? hh : -hh;
            const uint32_t j = k % 5u;
            if (j == 0u) p = mk(0.0f, k < 5u ? hh + r : -(hh + r), 0.0f);
            else if (j == 1u) p = mk(r, y, 0.0f);
            else if (j == 2u) p = mk(-r, y, 0.0f);
            else if (j == 3u) p = mk(0.0f, y, r);
            else p = mk(0.0f, y, -r);
            return true;
        }
        p = mk(k == 10u ? r : -r, 0.0f, 0.0f);  // shaft midpoints
        return true;
    }
    return false;
}

// glam Quat * Vec3 (scalar form): v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w).
RKF_DEV f3 quat_mul_vec(const float q[4], f3 v) {
    const float bx = q[0], by = q[1], bz = q[2], w = q[3];
    const float b2 = bx * bx + by * by + bz * bz;
    const float s = w * w - b2;
    const float d2 = (v.x * bx + v.y * by + v.z * bz) * 2.0f;
    const float cx = by * v.z - bz * v.y, cy = bz * v.x - bx * v.z, cz = bx * v.y - by * v.x;
    const float w2 = w * 2.0f;
    return mk((v.x * s + bx * d2) + cx * w2, (v.y * s + by * d2) + cy * w2, (v.z * s + bz * d2) + cz * w2);
}

// SdfQueryable::gradient via gradient_central_diff (sdf_collision.rs:44-56).
RKF_DEV f3 field_gradient(const SparseDev &S, f3 p) {
    BrickCache bc;
    bc_init(S, p, 3, bc);
    const float e = 0.01f;
    float dx = sp_sample_bc(S, bc, add(p, mk(e, 0.0f, 0.0f)), false) - sp_sample_bc(S, bc, sub(p, mk(e, 0.0f, 0.0f)), false);
    float dy = sp_sample_bc(S, bc, add(p, mk(0.0f, e, 0.0f)), false) - sp_sample_bc(S, bc, sub(p, mk(0.0f, e, 0.0f)), false);
    float dz = sp_sample_bc(S, bc, add(p, mk(0.0f, 0.0f, e)), false) - sp_sample_bc(S, bc, sub(p, mk(0.0f, 0.0f, e)), false);
    float rcp = 1.0f / __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    if (isfinite(rcp) && rcp > 0.0f) return mk(dx * rcp, dy * rcp, dz * rcp);
    return mk(0.0f, 0.0f, 0.0f);
}

}  // namespace rkf
