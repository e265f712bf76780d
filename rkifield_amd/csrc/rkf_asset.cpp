// rkf_asset.cpp — the .rkf v2 object asset (crates/rkf-core/src/asset_file.rs)
// read and written on the host, and streamed into resident sparse scenes.
//
// Layout (asset_file.rs:74-150): a 128-byte header ("RKF2", version 2, LOD
// count, material count, object AABB, analytical bound, 16 material ids),
// then one 40-byte entry per LOD (voxel size, brick count, brick-map dims,
// offset and sizes of its data block). Each block is the LZ4 block format
// with a 4-byte little-endian uncompressed-size prefix (lz4_flex
// compress_prepend_size, :458 / decompress_size_prepended, :579) of the
// dense brick map (dims product u32, bricks renumbered 0..n-1 in order of
// first appearance, :281-331) followed by the bricks (512 VoxelSample each).
// LODs are stored coarsest first (:388-395).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "host_internal.h"

namespace {

constexpr uint32_t kHeaderBytes = 128, kLodBytes = 40;

// LZ4 block decoder (public LZ4 block format): sequences of a token
// (literal length high nibble, match length - 4 low nibble, 15 = more bytes
// follow), literals, a 2-byte offset and the extra match-length bytes; the
// last sequence carries literals only. Every read and write is bounds-checked.
bool lz4_decode(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t cap, uint64_t &out) {
    uint64_t i = 0, o = 0;
    while (i < n) {
        const uint32_t token = src[i++];
        uint64_t lit = token >> 4;
        if (lit == 15u) {
            uint8_t b = 255;
            while (b == 255) {
                if (i >= n) return false;
                b = src[i++];
                lit += b;
            }
        }
        if (lit > n - i || lit > cap - o) return false;
        memcpy(dst + o, src + i, lit);
        i += lit;
        o += lit;
        if (i == n) break;
        if (n - i < 2) return false;
        const uint64_t off = (uint64_t)src[i] | ((uint64_t)src[i + 1] << 8);
        i += 2;
        if (off == 0 || off > o) return false;
        uint64_t len = (token & 15u) + 4u;
        if ((token & 15u) == 15u) {
            uint8_t b = 255;
            while (b == 255) {
                if (i >= n) return false;
                b = src[i++];
                len += b;
            }
        }
        if (len > cap - o) return false;
        for (uint64_t k = 0; k < len; k++) dst[o + k] = dst[o - off + k];  // overlap allowed
        o += len;
    }
    out = o;
    return true;
}

void put_len(std::vector<uint8_t> &v, uint64_t extra) {
    while (extra >= 255) {
        v.push_back(255);
        extra -= 255;
    }
    v.push_back((uint8_t)extra);
}

// Greedy LZ4 block encoder: 4-byte hash -> last position; a match needs 4
// equal bytes within 65535 back; the format's end rules (no match in the last
// 12 bytes, last 5 bytes literal) are kept.
void lz4_encode(const uint8_t *src, uint64_t n, std::vector<uint8_t> &out) {
    constexpr uint32_t kHashBits = 16;
    std::vector<int64_t> table(1u << kHashBits, -1);
    auto hash = [&](uint64_t p) {
        uint32_t v;
        memcpy(&v, src + p, 4);
        return (v * 2654435761u) >> (32 - kHashBits);
    };
    uint64_t anchor = 0, i = 0;
    const uint64_t limit = n >= 12 ? n - 12 : 0;
    auto emit = [&](uint64_t lit_end, uint64_t off, uint64_t mlen) {
        const uint64_t lit = lit_end - anchor;
        const bool has_match = mlen >= 4;
        const uint64_t ml = has_match ? mlen - 4 : 0;
        out.push_back((uint8_t)(((lit < 15 ? lit : 15) << 4) | (has_match ? (ml < 15 ? ml : 15) : 0)));
        if (lit >= 15) put_len(out, lit - 15);
        out.insert(out.end(), src + anchor, src + lit_end);
        if (has_match) {
            out.push_back((uint8_t)(off & 255));
            out.push_back((uint8_t)(off >> 8));
            if (ml >= 15) put_len(out, ml - 15);
        }
    };
    while (n >= 13 && i < limit) {
        const uint32_t h = hash(i);
        const int64_t cand = table[h];
        table[h] = (int64_t)i;
        if (cand >= 0 && i - (uint64_t)cand <= 65535 && memcmp(src + cand, src + i, 4) == 0) {
            uint64_t len = 4;
            const uint64_t max_len = n - 5 - i;  // keep the last 5 bytes literal
            while (len < max_len && src[cand + len] == src[i + len]) len++;
            emit(i, i - (uint64_t)cand, len);
            i += len;
            anchor = i;
            continue;
        }
        i++;
    }
    emit(n, 0, 0);  // trailing literals
}

uint32_t rd32(const uint8_t *p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

int parse(const uint8_t *f, uint64_t len, rkfm_rkf_header *h, std::vector<rkfm_rkf_lod> &lods) {
    if (!f || len < kHeaderBytes) return rkf_set_error(RKFM_ERR_FORMAT, "rkf: %llu bytes, header needs 128", (unsigned long long)len);
    memcpy(h, f, kHeaderBytes);
    if (memcmp(h->magic, "RKF2", 4) != 0) return rkf_set_error(RKFM_ERR_FORMAT, "rkf: invalid magic");
    if (h->version != 2u) return rkf_set_error(RKFM_ERR_FORMAT, "rkf: unsupported version %u", h->version);
    if ((uint64_t)h->lod_count * kLodBytes > len - kHeaderBytes)
        return rkf_set_error(RKFM_ERR_FORMAT, "rkf: %u LOD entries exceed the file", h->lod_count);
    lods.resize(h->lod_count);
    for (uint32_t i = 0; i < h->lod_count; i++) memcpy(&lods[i], f + kHeaderBytes + (uint64_t)i * kLodBytes, kLodBytes);
    return RKFM_OK;
}

}  // namespace

extern "C" {

int rkfm_rkf_read_header(const void *file, uint64_t len, rkfm_rkf_header *hdr, rkfm_rkf_lod *lods,
                         uint32_t max_lods) {
    if (!hdr) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkf_read_header: null header");
    std::vector<rkfm_rkf_lod> all;
    int rc = parse((const uint8_t *)file, len, hdr, all);
    if (rc) return rc;
    if (lods) memcpy(lods, all.data(), std::min<uint64_t>(all.size(), max_lods) * kLodBytes);
    return RKFM_OK;
}

int rkfm_rkf_decode_lod(const void *file, uint64_t len, uint32_t lod, uint32_t *map, rkf_voxel_sample *bricks) {
    rkfm_rkf_header h;
    std::vector<rkfm_rkf_lod> lods;
    const uint8_t *f = (const uint8_t *)file;
    int rc = parse(f, len, &h, lods);
    if (rc) return rc;
    if (lod >= h.lod_count)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "rkf: LOD index %u out of range (max %u)", lod, h.lod_count);
    const rkfm_rkf_lod &e = lods[lod];
    if (e.data_offset > len || e.compressed_size > len - e.data_offset || e.compressed_size < 4)
        return rkf_set_error(RKFM_ERR_FORMAT, "rkf: LOD %u data block outside the file", lod);
    const uint8_t *blk = f + e.data_offset;
    const uint64_t usz = rd32(blk);
    // An LZ4 sequence expands at most 255x (a 1-byte match-length code, a
    // 255-byte extension), so a larger size prefix is corrupt, not data; the
    // bound keeps a damaged file from sizing a multi-GiB buffer.
    if (usz > (uint64_t)(e.compressed_size - 4) * 255 + 16)
        return rkf_set_error(RKFM_ERR_FORMAT, "rkf: LOD %u: size prefix %llu exceeds what %u bytes can hold", lod,
                             (unsigned long long)usz, e.compressed_size - 4);
    uint64_t nmap = 0, need = 0;
    const bool wraps = __builtin_mul_overflow((uint64_t)e.brick_dims[0] * e.brick_dims[1], (uint64_t)e.brick_dims[2], &nmap) ||
                       __builtin_mul_overflow(nmap, (uint64_t)4, &need) ||
                       __builtin_add_overflow(need, (uint64_t)e.brick_count * 512 * sizeof(rkf_voxel_sample), &need);
    if (wraps || usz < need)
        return rkf_set_error(RKFM_ERR_FORMAT, "rkf: decompressed data is %llu bytes, expected at least %llu",
                             (unsigned long long)usz, (unsigned long long)need);
    std::vector<uint8_t> raw(usz);
    uint64_t got = 0;
    if (!lz4_decode(blk + 4, e.compressed_size - 4, raw.data(), usz, got) || got != usz)
        return rkf_set_error(RKFM_ERR_FORMAT, "rkf: LOD %u: corrupt LZ4 block", lod);
    if (map) memcpy(map, raw.data(), nmap * 4);
    if (bricks) memcpy(bricks, raw.data() + nmap * 4, need - nmap * 4);
    return RKFM_OK;
}

int rkfm_rkf_encode(const rkfm_rkf_header *hdr, const rkfm_rkf_lod_src *src, uint32_t n_lods, void *out,
                    uint64_t cap, uint64_t *len) {
    if (!hdr || !len || (n_lods && !src)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkf_encode: null argument");
    *len = 0;
    // coarsest first (descending voxel size, stable)
    std::vector<uint32_t> order(n_lods);
    for (uint32_t i = 0; i < n_lods; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return src[a].voxel_size > src[b].voxel_size; });
    std::vector<uint8_t> file(kHeaderBytes + (uint64_t)n_lods * kLodBytes, 0);
    rkfm_rkf_header h = *hdr;
    memcpy(h.magic, "RKF2", 4);
    h.version = 2;
    h.lod_count = n_lods;
    if (h.material_count > 16) h.material_count = 16;
    for (uint32_t k = h.material_count; k < 16; k++) h.material_ids[k] = 0;
    memset(h._reserved, 0, sizeof(h._reserved));
    memcpy(file.data(), &h, kHeaderBytes);
    for (uint32_t j = 0; j < n_lods; j++) {
        const rkfm_rkf_lod_src &L = src[order[j]];
        const uint64_t nmap = (uint64_t)L.dims[0] * L.dims[1] * L.dims[2];
        if (nmap && !L.map) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkf_encode: LOD %u map", order[j]);
        // renumber bricks by first appearance (asset_file.rs:281-313)
        std::unordered_map<uint32_t, uint32_t> local;
        std::vector<uint8_t> raw(nmap * 4);
        for (uint64_t i = 0; i < nmap; i++) {
            uint32_t s = L.map[i], v = RKF_EMPTY_SLOT;
            if (s != RKF_EMPTY_SLOT) {
                auto it = local.find(s);
                if (it == local.end()) it = local.emplace(s, (uint32_t)local.size()).first;
                v = it->second;
            }
            memcpy(raw.data() + i * 4, &v, 4);
        }
        if (local.size() != L.brick_count)
            return rkf_set_error(RKFM_ERR_LAYOUT, "rkf_encode: LOD %u map names %zu bricks, %u given", order[j],
                                 local.size(), L.brick_count);
        if (L.brick_count && !L.bricks) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkf_encode: LOD %u bricks", order[j]);
        const uint8_t *bb = (const uint8_t *)L.bricks;
        raw.insert(raw.end(), bb, bb + (uint64_t)L.brick_count * 512 * sizeof(rkf_voxel_sample));
        if (raw.size() > 0xFFFFFFFFull) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkf_encode: LOD %u too large", order[j]);
        std::vector<uint8_t> blk(4);
        const uint32_t usz = (uint32_t)raw.size();
        memcpy(blk.data(), &usz, 4);
        lz4_encode(raw.data(), raw.size(), blk);
        rkfm_rkf_lod e;
        memset(&e, 0, sizeof(e));
        e.voxel_size = L.voxel_size;
        e.brick_count = L.brick_count;
        memcpy(e.brick_dims, L.dims, 12);
        e.data_offset = file.size();
        e.compressed_size = (uint32_t)blk.size();
        e.uncompressed_size = usz;
        memcpy(file.data() + kHeaderBytes + (uint64_t)j * kLodBytes, &e, kLodBytes);
        file.insert(file.end(), blk.begin(), blk.end());
    }
    *len = file.size();
    if (!out) return RKFM_OK;
    if (cap < file.size())
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "rkf_encode: %llu bytes needed, %llu given",
                             (unsigned long long)file.size(), (unsigned long long)cap);
    memcpy(out, file.data(), file.size());
    return RKFM_OK;
}

int rkfm_sparse_from_rkf(rkfm_ctx *c, const void *file, uint64_t len, uint32_t lod, const float origin[3],
                         rkfm_sparse_scene **out) {
    if (!c || !out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_from_rkf: null argument");
    rkfm_rkf_header h;
    std::vector<rkfm_rkf_lod> lods;
    int rc = parse((const uint8_t *)file, len, &h, lods);
    if (rc) return rc;
    if (lod >= h.lod_count)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "rkf: LOD index %u out of range (max %u)", lod, h.lod_count);
    const rkfm_rkf_lod &e = lods[lod];
    // validate first: a damaged LOD table must not size the vectors below
    rc = rkfm_rkf_decode_lod(file, len, lod, nullptr, nullptr);
    if (rc) return rc;
    std::vector<uint32_t> map((uint64_t)e.brick_dims[0] * e.brick_dims[1] * e.brick_dims[2]);
    std::vector<rkf_voxel_sample> bricks((uint64_t)e.brick_count * 512);
    rc = rkfm_rkf_decode_lod(file, len, lod, map.data(), bricks.data());
    if (rc) return rc;
    float o[3];
    const float bws = e.voxel_size * 8.0f;
    for (int a = 0; a < 3; a++) o[a] = origin ? origin[a] : -((float)e.brick_dims[a] * bws * 0.5f);  // centred grid
    return rkfm_sparse_from_dense(c, map.data(), e.brick_dims, bricks.data(), e.brick_count, e.voxel_size, o, out);
}

}  // extern "C"
