// SPDX-License-Identifier: MIT
// World API: shape upload preparation, device buffers, stepping and
// result read-back in the reference's record layouts.
#include "pmh_world.hpp"
#include "pmh_warmstart.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <unordered_map>

namespace pmh {

static std::string g_last_error;
static std::mutex g_error_mutex;

PmhResult hip_result(hipError_t e, const char *what) {
    {
        std::lock_guard<std::mutex> lock(g_error_mutex);
        g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    }
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PMH_ERROR_OUT_OF_DEVICE_MEMORY;
    if (e == hipErrorInvalidValue) return PMH_ERROR_INVALID_ARGUMENT;
    return PMH_ERROR_DEVICE_LOST;
}
const std::string &last_error() { return g_last_error; }

// ---------------------------------------------------------------------
// Triangle-mesh upload preparation: validation, median-split BVH triangle
// order (<= 4 triangles per leaf), convex face planes with an exact-dedup
// tail, bounds, bounding sphere and box unit inertia, as
// SimulationRuntime::add_triangle_mesh (runtime_deformables.cu.inc:4-304,
// runtime_domains.cu.inc:1061-1108, src/convex_plane_dedup.hpp).
// ---------------------------------------------------------------------
static std::vector<std::array<float, 4>> convex_planes(const std::vector<V3> &v, const std::vector<uint32_t> &idx) {
    std::map<std::array<float, 3>, uint32_t> weld;
    std::vector<uint32_t> remap(v.size());
    V3 c{0.0f, 0.0f, 0.0f};
    for (size_t i = 0; i < v.size(); ++i) {
        auto ins = weld.emplace(std::array<float, 3>{v[i].x, v[i].y, v[i].z}, uint32_t(weld.size()));
        remap[i] = ins.first->second;
        if (ins.second) c = add(c, v[i]);
    }
    if (weld.size() < 4) return {};
    c = mul(c, 1.0f / float(weld.size()));
    std::unordered_map<uint64_t, uint32_t> edge_uses;
    for (size_t t = 0; t < idx.size(); t += 3)
        for (size_t e = 0; e < 3; ++e) {
            uint32_t a = remap[idx[t + e]], b = remap[idx[t + (e + 1) % 3]];
            ++edge_uses[(uint64_t(std::min(a, b)) << 32) | std::max(a, b)];
        }
    for (const auto &e : edge_uses) if (e.second != 2) return {};   // open or non-manifold
    float scale = 0.0f;
    for (const V3 &p : v) scale = std::max(scale, len(sub(p, c)));
    const float tol = std::max(1.0e-7f, scale * 1.0e-5f);
    std::vector<std::array<float, 4>> out;
    out.reserve(idx.size() / 3);
    for (size_t t = 0; t < idx.size(); t += 3) {
        const V3 a = v[idx[t]];
        V3 nrm = normalized_or(cross(sub(v[idx[t + 1]], a), sub(v[idx[t + 2]], a)), {0.0f, 0.0f, 0.0f});
        const float side = dot(nrm, sub(a, c));
        if (std::fabs(side) <= tol) return {};
        if (side < 0.0f) nrm = mul(nrm, -1.0f);
        const float off = dot(nrm, a);
        for (const V3 &p : v) if (dot(nrm, p) - off > tol) return {};   // not convex
        out.push_back({nrm.x, nrm.y, nrm.z, off});
    }
    return out;
}

// Outward normals certified per triangle at upload, as closed_shell_normals
// (runtime_domains.cu.inc:1111-1245): vertices welded by position are
// grouped into connected components; a convex component takes its face
// planes' normals, a closed orientable one the winding normal signed by its
// volume, an open volumetric one the side of each support plane (or a
// quarter-length two-sided marker on interior triangles).  Empty when no
// triangle is certified.
static std::vector<V3> shell_normals(const std::vector<V3> &v, const std::vector<uint32_t> &idx) {
    const uint32_t nv = uint32_t(v.size());
    std::map<std::array<float, 3>, uint32_t> weld;
    std::vector<uint32_t> rep(nv), parent(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        rep[i] = weld.emplace(std::array<float, 3>{v[i].x, v[i].y, v[i].z}, i).first->second;
        parent[i] = i;
    }
    auto root = [&](uint32_t i) {
        while (parent[i] != i) { parent[i] = parent[parent[i]]; i = parent[i]; }
        return i;
    };
    for (size_t t = 0; t < idx.size(); t += 3)
        for (size_t c = 1; c < 3; ++c) {
            const uint32_t r0 = root(rep[idx[t]]), rc = root(rep[idx[t + c]]);
            parent[rc] = r0;
        }
    const uint32_t tris = uint32_t(idx.size() / 3);
    std::map<uint32_t, std::vector<uint32_t>> parts;
    for (uint32_t t = 0; t < tris; ++t) parts[root(rep[idx[3 * t]])].push_back(t);
    std::vector<V3> out(tris, V3{0.0f, 0.0f, 0.0f});
    bool certified = false;
    for (const auto &part : parts) {
        const std::vector<uint32_t> &ts = part.second;
        std::map<uint32_t, uint32_t> local;
        std::vector<V3> pts;
        std::vector<uint32_t> faces;
        for (uint32_t t : ts)
            for (uint32_t c = 0; c < 3; ++c) {
                const uint32_t g = rep[idx[3 * t + c]];
                auto ins = local.emplace(g, uint32_t(pts.size()));
                if (ins.second) pts.push_back(v[g]);
                faces.push_back(ins.first->second);
            }
        const auto planes = convex_planes(pts, faces);
        if (!planes.empty()) {
            certified = true;
            for (size_t i = 0; i < ts.size(); ++i) out[ts[i]] = {planes[i][0], planes[i][1], planes[i][2]};
            continue;
        }
        std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, int>> edges;
        double volume = 0.0;
        for (size_t f = 0; f < faces.size(); f += 3) {
            for (size_t c = 0; c < 3; ++c) {
                const uint32_t e0 = faces[f + c], e1 = faces[f + (c + 1) % 3];
                auto &use = edges[{std::min(e0, e1), std::max(e0, e1)}];
                use.first += 1;
                use.second += e0 < e1 ? 1 : -1;
            }
            double d[3][3];
            for (size_t c = 0; c < 3; ++c) {
                const V3 q = pts[faces[f + c]];
                d[c][0] = double(q.x) - pts[0].x;
                d[c][1] = double(q.y) - pts[0].y;
                d[c][2] = double(q.z) - pts[0].z;
            }
            volume += d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) +
                      d[0][1] * (d[1][2] * d[2][0] - d[1][0] * d[2][2]) +
                      d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        }
        bool closed = volume != 0.0 && std::isfinite(volume);
        for (const auto &e : edges) closed = closed && e.second.first == 2u && e.second.second == 0;
        if (closed) {
            certified = true;
            for (size_t i = 0; i < ts.size(); ++i) {
                const V3 a = pts[faces[3 * i]], b = pts[faces[3 * i + 1]], c = pts[faces[3 * i + 2]];
                out[ts[i]] = mul(normalized_or(cross(sub(b, a), sub(c, a)), {0.0f, 0.0f, 0.0f}),
                                 volume > 0.0 ? 1.0f : -1.0f);
            }
            continue;
        }
        float scale = 0.0f;
        for (const V3 &q : pts) scale = std::max(scale, len(sub(q, pts[0])));
        const float tol = std::max(1.0e-7f, scale * 1.0e-5f);
        for (size_t i = 0; i < ts.size(); ++i) {
            const V3 a = pts[faces[3 * i]];
            const V3 nrm = normalized_or(cross(sub(pts[faces[3 * i + 1]], a), sub(pts[faces[3 * i + 2]], a)),
                                         {0.0f, 0.0f, 0.0f});
            float lo = 0.0f, hi = 0.0f;
            for (const V3 &q : pts) {
                const float d = dot(nrm, sub(q, a));
                lo = std::min(lo, d);
                hi = std::max(hi, d);
            }
            if (hi <= tol && lo < -tol) {
                out[ts[i]] = nrm;
            } else if (lo >= -tol && hi > tol) {
                out[ts[i]] = mul(nrm, -1.0f);
            } else {
                bool interior = true;
                for (size_t c = 0; c < 3; ++c) {
                    const uint32_t e0 = faces[3 * i + c], e1 = faces[3 * i + (c + 1) % 3];
                    const auto use = edges.at({std::min(e0, e1), std::max(e0, e1)});
                    interior = interior && use.first == 2u && use.second == 0;
                }
                if (!interior) continue;
                out[ts[i]] = mul(nrm, 0.25f);
            }
            certified = true;
        }
    }
    return certified ? out : std::vector<V3>{};
}

static PmhResult prepare_mesh(const float *vp, uint32_t nv, const uint32_t *ip, uint32_t ni, HostShape &sh) {
    if (!vp || !ip || nv < 3 || ni < 3 || ni % 3) return PMH_ERROR_INVALID_ARGUMENT;
    std::vector<V3> v(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        v[i] = {vp[3 * i], vp[3 * i + 1], vp[3 * i + 2]};
        if (!std::isfinite(v[i].x) || !std::isfinite(v[i].y) || !std::isfinite(v[i].z))
            return PMH_ERROR_INVALID_ARGUMENT;
    }
    std::vector<uint32_t> in(ip, ip + ni);
    V3 mn = v[0], mx = v[0];
    for (const V3 &p : v) { mn = cmin(mn, p); mx = cmax(mx, p); }
    for (uint32_t t = 0; t < ni; t += 3) {
        if (in[t] >= nv || in[t + 1] >= nv || in[t + 2] >= nv) return PMH_ERROR_INVALID_ARGUMENT;
        if (len2(cross(sub(v[in[t + 1]], v[in[t]]), sub(v[in[t + 2]], v[in[t]]))) <= kEpsilon * kEpsilon)
            return PMH_ERROR_INVALID_ARGUMENT;
    }
    const uint32_t tris = ni / 3;
    std::vector<uint32_t> order(tris);
    std::iota(order.begin(), order.end(), 0u);
    auto centroid = [&](uint32_t t) {
        return mul(add(add(v[in[3 * t]], v[in[3 * t + 1]]), v[in[3 * t + 2]]), 1.0f / 3.0f);
    };
    sh.nodes.clear();
    sh.nodes.reserve(2 * tris);
    std::function<uint32_t(uint32_t, uint32_t)> split = [&](uint32_t b, uint32_t e) -> uint32_t {
        V3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t k = b; k < e; ++k)
            for (uint32_t c = 0; c < 3; ++c) { lo = cmin(lo, v[in[3 * order[k] + c]]); hi = cmax(hi, v[in[3 * order[k] + c]]); }
        const uint32_t self = uint32_t(sh.nodes.size());
        BvhDev node{};
        node.mn[0] = lo.x; node.mn[1] = lo.y; node.mn[2] = lo.z;
        node.mx[0] = hi.x; node.mx[1] = hi.y; node.mx[2] = hi.z;
        sh.nodes.push_back(node);
        if (e - b <= 4) {
            sh.nodes[self].first = b;
            sh.nodes[self].count = e - b;
            return self;
        }
        const V3 ext = sub(hi, lo);
        const int axis = ext.x >= ext.y && ext.x >= ext.z ? 0 : (ext.y >= ext.z ? 1 : 2);
        std::stable_sort(order.begin() + b, order.begin() + e, [&](uint32_t x, uint32_t y) {
            const float fx = comp(centroid(x), axis), fy = comp(centroid(y), axis);
            return fx < fy || (fx == fy && x < y);
        });
        const uint32_t mid = b + (e - b) / 2;
        const uint32_t l = split(b, mid);
        const uint32_t r = split(mid, e);
        sh.nodes[self].left = l;
        sh.nodes[self].right = r;
        return self;
    };
    split(0, tris);
    sh.leaves.clear();
    for (uint32_t k = 0; k < sh.nodes.size(); ++k)
        if (sh.nodes[k].count != 0u) sh.leaves.push_back(k);
    sh.indices.resize(ni);
    for (uint32_t t = 0; t < tris; ++t)
        for (uint32_t c = 0; c < 3; ++c) sh.indices[3 * t + c] = in[3 * order[t] + c];
    sh.vertices = v;
    auto planes = convex_planes(v, sh.indices);
    sh.shell = shell_normals(v, sh.indices);
    sh.planes.clear();
    sh.unique_planes = 0;
    if (!planes.empty()) {
        std::map<std::array<float, 4>, bool> seen;
        std::vector<std::array<float, 4>> uniq;
        for (const auto &p : planes) if (seen.emplace(p, true).second) uniq.push_back(p);
        for (const auto &p : planes) sh.planes.insert(sh.planes.end(), p.begin(), p.end());
        for (const auto &p : uniq) sh.planes.insert(sh.planes.end(), p.begin(), p.end());
        sh.unique_planes = uint32_t(uniq.size());
    }
    ShapeDev &d = sh.info;
    std::memset(&d, 0, sizeof d);
    d.kind = kShapeMesh;
    d.vtx_count = nv;
    d.idx_count = ni;
    d.has_planes = planes.empty() ? 0u : 1u;
    d.unique_planes = sh.unique_planes;
    d.convex_face = d.has_planes && ni <= 96u;
    d.mn[0] = mn.x; d.mn[1] = mn.y; d.mn[2] = mn.z;
    d.mx[0] = mx.x; d.mx[1] = mx.y; d.mx[2] = mx.z;
    const V3 c = mul(add(mn, mx), 0.5f);
    d.center[0] = c.x; d.center[1] = c.y; d.center[2] = c.z;
    float r2 = 0.0f;
    for (const V3 &p : v) r2 = std::fmax(r2, len2(sub(p, c)));
    d.bradius = std::sqrt(r2);
    const V3 h = mul(sub(mx, mn), 0.5f);
    sh.unit_inertia[0] = std::fmax((h.y * h.y + h.z * h.z) / 3.0f, kEpsilon);
    sh.unit_inertia[1] = std::fmax((h.x * h.x + h.z * h.z) / 3.0f, kEpsilon);
    sh.unit_inertia[2] = std::fmax((h.x * h.x + h.y * h.y) / 3.0f, kEpsilon);
    // A centred axis-aligned box also gets an analytic description so
    // round shapes can collide with it in closed form.
    bool box = d.has_planes && nv == 8 && c.x == 0.0f && c.y == 0.0f && c.z == 0.0f;
    for (const V3 &p : v)
        box = box && std::fabs(p.x) == h.x && std::fabs(p.y) == h.y && std::fabs(p.z) == h.z;
    d.is_box = box ? 1u : 0u;
    d.box_half[0] = h.x; d.box_half[1] = h.y; d.box_half[2] = h.z;
    return PMH_SUCCESS;
}

static uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

PmhResult World::ensure_device() {
    if (hipSetDevice(device) != hipSuccess) return PMH_ERROR_DEVICE_LOST;
    if (!stream) PMH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (!events[0])
        for (auto &e : events) PMH_CHECK(hipEventCreate(&e));
    if (!solve_timing.start) {
        PMH_CHECK(hipEventCreate(&solve_timing.start));
        PMH_CHECK(hipEventCreate(&solve_timing.stop));
    }
    PmhResult r;
    if (shapes_dirty) {
        std::vector<ShapeDev> info;
        std::vector<V3> verts;
        std::vector<uint32_t> idx;
        std::vector<float4> planes, shell;
        std::vector<BvhDev> nodes;
        std::vector<uint32_t> leaves;
        for (HostShape &s : shapes) {
            ShapeDev d = s.info;
            d.vtx_off = uint32_t(verts.size());
            d.idx_off = uint32_t(idx.size());
            d.plane_off = uint32_t(planes.size());
            d.node_off = uint32_t(nodes.size());
            d.node_count = uint32_t(s.nodes.size());
            d.leaf_off = uint32_t(leaves.size());
            d.leaf_count = uint32_t(s.leaves.size());
            d.shell_off = uint32_t(shell.size());
            d.has_shell = s.shell.empty() ? 0u : 1u;
            verts.insert(verts.end(), s.vertices.begin(), s.vertices.end());
            idx.insert(idx.end(), s.indices.begin(), s.indices.end());
            nodes.insert(nodes.end(), s.nodes.begin(), s.nodes.end());
            leaves.insert(leaves.end(), s.leaves.begin(), s.leaves.end());
            for (const V3 &o : s.shell) shell.push_back(make_float4(o.x, o.y, o.z, 0.0f));
            for (size_t k = 0; k < s.planes.size(); k += 4)
                planes.push_back(make_float4(s.planes[k], s.planes[k + 1], s.planes[k + 2], s.planes[k + 3]));
            info.push_back(d);
        }
        if ((r = shapes_dev.reserve(info.size())) || (r = verts_dev.reserve(verts.size())) ||
            (r = idx_dev.reserve(idx.size())) || (r = planes_dev.reserve(planes.size())) ||
            (r = nodes_dev.reserve(nodes.size())) || (r = leaves_dev.reserve(leaves.size())) ||
            (r = shell_dev.reserve(shell.size())))
            return r;
        if (!nodes.empty()) PMH_CHECK(hipMemcpy(nodes_dev.p, nodes.data(), nodes.size() * sizeof(BvhDev), hipMemcpyHostToDevice));
        if (!leaves.empty()) PMH_CHECK(hipMemcpy(leaves_dev.p, leaves.data(), leaves.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!shell.empty()) PMH_CHECK(hipMemcpy(shell_dev.p, shell.data(), shell.size() * sizeof(float4), hipMemcpyHostToDevice));
        if (!info.empty()) PMH_CHECK(hipMemcpy(shapes_dev.p, info.data(), info.size() * sizeof(ShapeDev), hipMemcpyHostToDevice));
        if (!verts.empty()) PMH_CHECK(hipMemcpy(verts_dev.p, verts.data(), verts.size() * sizeof(V3), hipMemcpyHostToDevice));
        if (!idx.empty()) PMH_CHECK(hipMemcpy(idx_dev.p, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!planes.empty()) PMH_CHECK(hipMemcpy(planes_dev.p, planes.data(), planes.size() * sizeof(float4), hipMemcpyHostToDevice));
        shapes_dirty = false;
    }
    if (bodies_dirty) {
        const uint32_t n = body_count;
        if (n > 0x7fffffffu) return PMH_ERROR_CAPACITY_EXCEEDED;
        // Per-body candidate lists hold up to pairs_per_body entries; the
        // pair and point pools are sized by their realistic totals.
        const uint32_t pair_cap = std::max<uint32_t>(256, n * pairs_per_body);
        const uint32_t point_cap = std::max<uint32_t>(1024, n * points_per_body);
        table_size = next_pow2(std::max<uint32_t>(1024, 2 * n));
        // Grid cell: 1.25 x the 95th percentile of the dynamic bodies' proxy
        // width at rest (bounding sphere, margin, epsilon and the bounded
        // solver correction on both sides); wider proxies go to the large list.
        std::vector<float> widths;
        for (uint32_t i = 0; i < n; ++i) {
            const PmhBodyParameters &b = host_params[i];
            if (b.motion != PMH_MOTION_DYNAMIC) continue;
            const ShapeDev &d = shapes[b.mesh_index].info;
            const V3 e = sub(V3{d.mx[0], d.mx[1], d.mx[2]}, V3{d.mn[0], d.mn[1], d.mn[2]});
            const float corr = std::min(0.5f, 0.25f * std::max(0.0f, std::min(e.x, std::min(e.y, e.z))));
            const float c = std::sqrt(d.center[0] * d.center[0] + d.center[1] * d.center[1] + d.center[2] * d.center[2]);
            widths.push_back(2.0f * (c + d.bradius + b.collision_margin + 1.0e-5f + std::min(1.0f, corr)));
        }
        if (!widths.empty()) {
            const size_t k = std::min(widths.size() - 1, size_t(0.95 * double(widths.size())));
            std::nth_element(widths.begin(), widths.begin() + k, widths.end());
            grid_cell = 1.25f * widths[k];
        }
        const uint64_t cap_n = body_capacity ? body_capacity : n;
        cache_capacity = uint32_t(std::min<uint64_t>(cap_n * (cap_n - 1) / 2, std::max<uint64_t>(4096, 8 * cap_n)));
        // resource_plan.hpp:151-164, geometry_storage.hpp:164-166
        leaf_slots = cache_capacity;
        leaf_limit = cap_n <= 8 ? 4096u : 512u;
        cache_capacity = std::max<uint32_t>(1, cache_capacity);
        if ((r = params.reserve(n)) || (r = state_cur.reserve(n)) || (r = state_prev.reserve(n)) ||
            (r = state_next.reserve(n)) || (r = proxy_mn.reserve(n)) || (r = proxy_mx.reserve(n)) ||
            (r = cell_of.reserve(n)) || (r = bucket_count.reserve(table_size + 1)) ||
            (r = bucket_start.reserve(table_size + 1)) || (r = entry_cell.reserve(n)) ||
            (r = entry_mn.reserve(n)) || (r = entry_mx.reserve(n)) || (r = large.reserve(n)) ||
            (r = cand_count.reserve(n + 1)) || (r = cand_offset.reserve(n + 1)) ||
            (r = cache_n.reserve(2)) || (r = inputs_dev.reserve(4 * size_t(n))) ||
            (r = targets_dev.reserve(n)) || (r = target_flags_dev.reserve(n)))
            return r;
        PMH_CHECK(hipMemset(inputs_dev.p, 0, sizeof(float4) * 4 * size_t(n)));
        PMH_CHECK(hipMemset(target_flags_dev.p, 0, sizeof(uint32_t) * n));
        for (int k = 0; k < 2; ++k)
            if ((r = cache_code[k].reserve(cache_capacity)) || (r = cache_count[k].reserve(cache_capacity)) ||
                (r = cache_data[k].reserve(size_t(cache_capacity) * 8 * kCacheFloat4PerPoint)))
                return r;
        const bool grew = solver.n != n;
        if ((r = solver.reserve(n, pair_cap, point_cap))) return r;
        if (grew) PMH_CHECK(hipMemset(solver.prev_vel.p, 0, sizeof(float4) * n));
        PMH_CHECK(hipMemcpy(params.p, host_params.data(), sizeof(PmhBodyParameters) * n, hipMemcpyHostToDevice));
        PMH_CHECK(hipMemcpy(state_cur.p, host_states.data(), sizeof(PmhRigidBodyState) * n, hipMemcpyHostToDevice));
        PMH_CHECK(hipMemset(cache_n.p, 0, sizeof(uint32_t) * 2));
        if (scene_first.size() > 1) {
            const uint32_t ns = uint32_t(scene_first.size());
            std::vector<uint32_t> of(n), cnt(ns);
            for (uint32_t sc = 0; sc < ns; ++sc) {
                const uint32_t b = scene_first[sc], e = sc + 1 < ns ? scene_first[sc + 1] : n;
                cnt[sc] = e - b;
                for (uint32_t i = b; i < e; ++i) of[i] = sc;
            }
            if ((r = scene_of.reserve(n)) || (r = scene_count.reserve(ns)) || (r = scene_pair_start.reserve(ns)) ||
                (r = scene_first_dev.reserve(ns)) || (r = scene_large_count.reserve(ns + 1)) ||
                (r = scene_large_start.reserve(ns + 1)) || (r = scene_large.reserve(n)))
                return r;
            PMH_CHECK(hipMemcpy(scene_of.p, of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
            PMH_CHECK(hipMemcpy(scene_count.p, cnt.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
            PMH_CHECK(hipMemcpy(scene_first_dev.p, scene_first.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
        }
        bodies_dirty = false;
        joints_dirty = true;   // exclusion lists are sized by the body count
    }
    if (joints_dirty && (r = upload_joints()) != PMH_SUCCESS) return r;
    if ((r = reserve_warm()) != PMH_SUCCESS) return r;
    if (inputs_dirty) {
        // The step consumes and clears them (pm_clear_rigid_inputs).
        PMH_CHECK(hipMemcpyAsync(inputs_dev.p, host_inputs.data(), sizeof(float4) * host_inputs.size(),
                                 hipMemcpyHostToDevice, stream));
        PMH_CHECK(hipMemcpyAsync(targets_dev.p, host_targets.data(), sizeof(PmhRigidBodyState) * host_targets.size(),
                                 hipMemcpyHostToDevice, stream));
        PMH_CHECK(hipMemcpyAsync(target_flags_dev.p, host_target_flags.data(),
                                 sizeof(uint32_t) * host_target_flags.size(), hipMemcpyHostToDevice, stream));
        PMH_CHECK(hipStreamSynchronize(stream));
        std::fill(host_inputs.begin(), host_inputs.end(), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        std::fill(host_target_flags.begin(), host_target_flags.end(), 0u);
        inputs_dirty = false;
    }
    return PMH_SUCCESS;
}

// Joint records and their collision exclusions (geometry.slang:494-511,
// 3108-3136).  Body ids resolve to {dense index, generation 1}; a joint
// excludes its pair while it is alive, enabled, unbroken and has
// disable_collisions, and a fixed one also welds its bodies together.
PmhResult World::upload_joints() {
    const uint32_t nj = uint32_t(host_joints.size());
    const uint32_t n = body_count;
    PmhResult r;
    if ((r = joints_dev.reserve(std::max<uint32_t>(1, nj))) || (r = solver.reserve_joints(nj)) ||
        (r = weld.reserve(std::max<uint32_t>(1, n))) || (r = excl_start.reserve(n + 1)))
        return r;
    if (nj) PMH_CHECK(hipMemcpy(joints_dev.p, host_joints.data(), sizeof(PmhRigidConstraint) * nj,
                                hipMemcpyHostToDevice));
    std::vector<uint32_t> start(n + 1, 0u);
    std::vector<uint2> list;
    weld_joints = 0;
    bool hinges = false;
    auto dense = [&](PmhBodyId id) { return id.index < n && id.generation == 1u ? id.index : kInvalid; };
    for (const PmhRigidConstraint &j : host_joints) {
        if (j.options.type == kJointHinge) hinges = true;
        const uint32_t a = dense(j.options.body_a), b = dense(j.options.body_b);
        if (j.alive == 0u || j.options.enabled == 0u || j.options.disable_collisions == 0u) continue;
        if (a == kInvalid || b == kInvalid) continue;
        start[a]++;
        start[b]++;
        if (j.options.type == kJointFixed) weld_joints++;
    }
    uint32_t total = 0;
    for (uint32_t i = 0; i <= n; ++i) { const uint32_t c = start[i]; start[i] = total; total += c; }
    list.resize(std::max<uint32_t>(1, total));
    std::vector<uint32_t> cursor(start.begin(), start.end());
    for (uint32_t k = 0; k < nj; ++k) {
        const PmhRigidConstraint &j = host_joints[k];
        const uint32_t a = dense(j.options.body_a), b = dense(j.options.body_b);
        if (j.alive == 0u || j.options.enabled == 0u || j.options.disable_collisions == 0u) continue;
        if (a == kInvalid || b == kInvalid) continue;
        list[cursor[a]++] = make_uint2(b, k);
        list[cursor[b]++] = make_uint2(a, k);
    }
    // Candidate joint lists and the hinge projection schedule.
    JointLists jl;
    build_joint_lists(host_joints.data(), nj, n, jl);
    if ((r = jlist_start.reserve(n + 1)) || (r = jlist.reserve(jl.list.size())) ||
        (r = jheavy.reserve(std::max<size_t>(1, jl.heavy.size()))))
        return r;
    PMH_CHECK(hipMemcpy(jlist_start.p, jl.start.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    PMH_CHECK(hipMemcpy(jlist.p, jl.list.data(), sizeof(uint32_t) * jl.list.size(), hipMemcpyHostToDevice));
    if (!jl.heavy.empty())
        PMH_CHECK(hipMemcpy(jheavy.p, jl.heavy.data(), sizeof(uint32_t) * jl.heavy.size(), hipMemcpyHostToDevice));
    solver.jlist_start = jlist_start.p;
    solver.jlist = jlist.p;
    solver.jheavy = jheavy.p;
    solver.nheavy = uint32_t(jl.heavy.size());
    if ((r = solver.jadj.reserve(std::max<size_t>(1, jl.list.size())))) return r;
    excl_count = total;
    if ((r = excl_list.reserve(list.size()))) return r;
    PMH_CHECK(hipMemcpy(excl_start.p, start.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    PMH_CHECK(hipMemcpy(excl_list.p, list.data(), sizeof(uint2) * list.size(), hipMemcpyHostToDevice));
    if ((r = hinge_list.reserve(std::max<size_t>(1, jl.hinges.size())))) return r;
    if (!jl.hinges.empty())
        PMH_CHECK(hipMemcpy(hinge_list.p, jl.hinges.data(), sizeof(uint32_t) * jl.hinges.size(), hipMemcpyHostToDevice));
    solver.hinge_list = hinge_list.p;
    solver.hinge_level_start = jl.level_start;
    if ((r = hinge_levels.reserve(solver.hinge_level_start.size()))) return r;
    PMH_CHECK(hipMemcpy(hinge_levels.p, solver.hinge_level_start.data(),
                        sizeof(uint32_t) * solver.hinge_level_start.size(), hipMemcpyHostToDevice));
    solver.hinge_level_dev = hinge_levels.p;
    solver.joints = joints_dev.p;
    solver.nj = nj;
    solver.hinges = hinges && !jl.hinges.empty();
    joints_dirty = false;
    joints_on_device = nj != 0;
    return PMH_SUCCESS;
}

// Cold-point buffers of the learned warm start, when weights or the
// contact log are set: up to 2^21 cold points a step.
PmhResult World::reserve_warm() {
    solver.mlp = warm_enabled ? mlp_dev.p : nullptr;
    solver.mlp_policy = warm_policy;
    solver.log = log_cap ? log_dev.p : nullptr;
    solver.log_cap = log_cap;
    if (!warm_enabled && !log_cap) return PMH_SUCCESS;
    const uint32_t cap = std::max<uint32_t>(64, std::min<uint32_t>(solver.point_cap, 1u << 21));
    PmhResult r;
    if ((r = solver.cold_list.reserve(cap)) || (r = solver.cold_feat.reserve(size_t(cap) * kMlpIn)) ||
        (r = solver.cold_scale.reserve(2 * size_t(cap))))
        return r;
    return PMH_SUCCESS;
}

// One step with the learned warm start in to-residual mode: when a phase
// ends above the tolerance, the step is taken again from the same inputs
// with the plain warm start (the step reads its pre-step states and never
// writes them; the velocity history, the one-step inputs and the joints it
// does write are restored from copies).
PmhResult World::step_with_fallback(uint32_t k) {
    const uint32_t n = body_count, nj = solver.nj;
    PmhResult r;
    if ((r = snap_vel.reserve(n)) || (r = snap_inputs.reserve(4 * size_t(n))) || (r = snap_flags.reserve(n)) ||
        (nj && (r = snap_joints.reserve(nj))))
        return r;
    PMH_CHECK(hipMemcpyAsync(snap_vel.p, solver.prev_vel.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
    PMH_CHECK(hipMemcpyAsync(snap_inputs.p, inputs_dev.p, sizeof(float4) * 4 * n, hipMemcpyDeviceToDevice, stream));
    PMH_CHECK(hipMemcpyAsync(snap_flags.p, target_flags_dev.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
    if (nj)
        PMH_CHECK(hipMemcpyAsync(snap_joints.p, joints_dev.p, sizeof(PmhRigidConstraint) * nj, hipMemcpyDeviceToDevice,
                                 stream));
    const int parity = cache_parity;
    const uint32_t sub = substep;
    if ((r = step_once(k, true)) != PMH_SUCCESS) return r;
    std::unique_ptr<ResidualLog> held(new ResidualLog);
    ResidualLog &log = *held;
    PMH_CHECK(hipMemcpyAsync(&log, solver.residual.p, sizeof log, hipMemcpyDeviceToHost, stream));
    PMH_CHECK(hipStreamSynchronize(stream));
    bool failed = false;
    for (uint32_t ph = 0; ph < 2; ++ph) {
        const uint32_t ex = log.executed[ph];
        if (ex == 0 || ex > kResidualSlots) continue;
        const uint32_t slot = ph * kResidualSlots + ex - 1;
        const ResidualTotal t = residual_fold(log, slot);
        failed = failed || residual_rms(t.sum, t.count) > residual_tolerance;
    }
    ++warm_steps;
    if (!failed) return PMH_SUCCESS;
    ++warm_fallbacks;
    std::swap(state_cur.p, state_next.p);
    cache_parity = parity;
    substep = sub;
    PMH_CHECK(hipMemcpyAsync(solver.prev_vel.p, snap_vel.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
    PMH_CHECK(hipMemcpyAsync(inputs_dev.p, snap_inputs.p, sizeof(float4) * 4 * n, hipMemcpyDeviceToDevice, stream));
    PMH_CHECK(hipMemcpyAsync(target_flags_dev.p, snap_flags.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
    if (nj)
        PMH_CHECK(hipMemcpyAsync(joints_dev.p, snap_joints.p, sizeof(PmhRigidConstraint) * nj, hipMemcpyDeviceToDevice,
                                 stream));
    return step_once(k, false);
}

// The device records (state, rows, duals) back into the host copy.
PmhResult World::pull_joints() {
    if (!joints_on_device) return PMH_SUCCESS;
    PMH_CHECK(hipStreamSynchronize(stream));
    PMH_CHECK(hipMemcpy(host_joints.data(), joints_dev.p, sizeof(PmhRigidConstraint) * solver.nj,
                        hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}

void World::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    shapes_dev.release(); verts_dev.release(); idx_dev.release(); planes_dev.release();
    nodes_dev.release(); leaves_dev.release(); shell_dev.release();
    inputs_dev.release(); targets_dev.release(); target_flags_dev.release();
    joints_dev.release(); weld.release(); excl_start.release(); excl_list.release();
    jlist_start.release(); jlist.release(); jheavy.release(); hinge_frames.release(); hinge_ref.release(); hinge_list.release(); hinge_levels.release();
    scene_of.release(); scene_count.release(); scene_pair_start.release(); scene_first_dev.release();
    scene_large_count.release(); scene_large_start.release(); scene_large.release();
    mlp_dev.release(); log_dev.release(); snap_vel.release(); snap_inputs.release(); snap_flags.release();
    snap_joints.release();
    params.release(); state_cur.release(); state_prev.release(); state_next.release();
    proxy_mn.release(); proxy_mx.release(); cell_of.release(); bucket_count.release();
    bucket_start.release(); large.release();
    entry_cell.release(); entry_mn.release(); entry_mx.release();
    cand_count.release(); cand_offset.release();
    for (int k = 0; k < 2; ++k) { cache_code[k].release(); cache_count[k].release(); cache_data[k].release(); }
    cache_n.release();
    solver.release();
    for (auto &e : events) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (auto &e : solve_events) if (e) (void)hipEventDestroy(e);
    solve_events.clear();
    kernel_timer.release();
    if (solve_timing.start) (void)hipEventDestroy(solve_timing.start);
    if (solve_timing.stop) (void)hipEventDestroy(solve_timing.stop);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

}  // namespace pmh

using pmh::World;

static World *W(PmhWorld w) { return reinterpret_cast<World *>(w); }

template <class T>
static PmhResult fetch(std::vector<T> &out, const T *src, size_t count) {
    out.resize(count);
    if (count) PMH_CHECK(hipMemcpy(out.data(), src, sizeof(T) * count, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}


extern "C" {

// The realtime contract (rigid_realtime.hpp:16-26, checked by
// pmCmdStepSimulation at src/api/resources.cpp:1240-1241): a 1/60 s step,
// within one float ulp either side, and one substep.
static bool realtime_step_valid(float timestep, uint32_t substeps) {
    const float dt = 1.0f / 60.0f;
    return std::isfinite(timestep) && timestep >= std::nextafter(dt, 0.0f) &&
           timestep <= std::nextafter(dt, INFINITY) && substeps <= 1u;
}

PmhResult pmhCreateWorld(const PmhWorldDesc *desc, PmhWorld *out) {
    if (!desc || !out || !(desc->timestep > 0.0f) || !std::isfinite(desc->timestep) || desc->profile_version > 1)
        return PMH_ERROR_INVALID_ARGUMENT;
    if (desc->profile_version == 1 && !realtime_step_valid(desc->timestep, desc->substeps))
        return PMH_ERROR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= int(desc->device)) return PMH_ERROR_FEATURE_NOT_PRESENT;
    World *w = new (std::nothrow) World();
    if (!w) return PMH_ERROR_OUT_OF_HOST_MEMORY;
    w->gravity = {desc->gravity[0], desc->gravity[1], desc->gravity[2]};
    if (desc->substeps > 64) {
        delete w;
        return PMH_ERROR_INVALID_ARGUMENT;
    }
    w->substeps = desc->substeps ? desc->substeps : 1u;
    w->dt = w->substeps == 1u ? desc->timestep : desc->timestep / float(w->substeps);
    w->profile = desc->profile_version;
    w->base_iterations = desc->base_iterations ? desc->base_iterations : 10;
    w->requested_iterations = desc->requested_iterations;
    w->body_capacity = desc->body_capacity;
    w->pairs_per_body = desc->pairs_per_body ? desc->pairs_per_body : 96;
    w->points_per_body = desc->points_per_body ? desc->points_per_body : 256;
    w->cooperative = (desc->flags & PMH_WORLD_COOPERATIVE_SOLVE) != 0;
    w->per_color_sweep = (desc->flags & PMH_WORLD_PER_COLOR_SWEEP) != 0;
    if (desc->solve_lanes != 0 && desc->solve_lanes != 8 && desc->solve_lanes != 16 && desc->solve_lanes != 32 &&
        desc->solve_lanes != 64) {
        delete w;
        return PMH_ERROR_INVALID_ARGUMENT;
    }
    w->solve_lanes = desc->solve_lanes ? desc->solve_lanes : pmh::kDefaultLanes;
    if (!(desc->residual_tolerance >= 0.0f) || !std::isfinite(desc->residual_tolerance)) {
        delete w;
        return PMH_ERROR_INVALID_ARGUMENT;
    }
    w->residual_tolerance = desc->residual_tolerance;
    w->kernel_timer.stride = desc->kernel_timing_stride;
    w->device = int(desc->device);
    w->timing_enabled = (desc->flags & PMH_WORLD_PHASE_TIMING) != 0;
    *out = reinterpret_cast<PmhWorld>(w);
    return PMH_SUCCESS;
}

void pmhDestroyWorld(PmhWorld world) {
    if (!world) return;
    W(world)->release();
    delete W(world);
}

PmhResult pmhWorldAddMesh(PmhWorld world, const float *vertices, uint32_t vertex_count, const uint32_t *indices,
                          uint32_t index_count, uint32_t *shape) {
    if (!world || !shape) return PMH_ERROR_INVALID_ARGUMENT;
    pmh::HostShape s;
    PmhResult r = pmh::prepare_mesh(vertices, vertex_count, indices, index_count, s);
    if (r != PMH_SUCCESS) return r;
    World *w = W(world);
    w->shapes.push_back(std::move(s));
    w->shapes_dirty = true;
    *shape = uint32_t(w->shapes.size() - 1);
    return PMH_SUCCESS;
}

static PmhResult add_round(World *w, uint32_t kind, float r, float h, uint32_t *shape) {
    pmh::HostShape s;
    pmh::ShapeDev &d = s.info;
    std::memset(&d, 0, sizeof d);
    d.kind = kind;
    d.sphere_radius = r;
    d.half_length = h;
    d.mn[0] = -r; d.mn[1] = -(h + r); d.mn[2] = -r;
    d.mx[0] = r; d.mx[1] = h + r; d.mx[2] = r;
    d.bradius = h + r;
    const float hy = h + r;
    if (kind == pmh::kShapeSphere) {
        s.unit_inertia[0] = s.unit_inertia[1] = s.unit_inertia[2] = 0.4f * r * r;
    } else {
        s.unit_inertia[0] = (hy * hy + r * r) / 3.0f;
        s.unit_inertia[1] = (r * r + r * r) / 3.0f;
        s.unit_inertia[2] = (r * r + hy * hy) / 3.0f;
    }
    w->shapes.push_back(s);
    w->shapes_dirty = true;
    *shape = uint32_t(w->shapes.size() - 1);
    return PMH_SUCCESS;
}

PmhResult pmhWorldAddSphere(PmhWorld world, float radius, uint32_t *shape) {
    if (!world || !shape || !(radius > 0.0f) || !std::isfinite(radius)) return PMH_ERROR_INVALID_ARGUMENT;
    return add_round(W(world), pmh::kShapeSphere, radius, 0.0f, shape);
}

PmhResult pmhWorldAddCapsule(PmhWorld world, float radius, float half_length, uint32_t *shape) {
    if (!world || !shape || !(radius > 0.0f) || !(half_length >= 0.0f) || !std::isfinite(radius + half_length))
        return PMH_ERROR_INVALID_ARGUMENT;
    return add_round(W(world), pmh::kShapeCapsule, radius, half_length, shape);
}

PmhResult pmhWorldGetShapeInfo(PmhWorld world, uint32_t shape, float out[16]) {
    if (!world || !out || shape >= W(world)->shapes.size()) return PMH_ERROR_INVALID_ARGUMENT;
    const pmh::HostShape &s = W(world)->shapes[shape];
    const pmh::ShapeDev &d = s.info;
    const float v[16] = {d.mn[0], d.mn[1], d.mn[2], d.mx[0], d.mx[1], d.mx[2], d.center[0], d.center[1],
                         d.center[2], d.bradius, s.unit_inertia[0], s.unit_inertia[1], s.unit_inertia[2],
                         float(d.kind), float(d.idx_count), float(s.unique_planes)};
    std::memcpy(out, v, sizeof v);
    return PMH_SUCCESS;
}

PmhResult pmhWorldAddBodies(PmhWorld world, uint32_t count, const PmhBodyParameters *parameters,
                            const PmhRigidBodyState *states) {
    if (!world || (count && (!parameters || !states))) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    for (uint32_t i = 0; i < count; ++i)
        if (parameters[i].mesh_index >= w->shapes.size() || parameters[i].motion > PMH_MOTION_DYNAMIC)
            return PMH_ERROR_INVALID_ARGUMENT;
    if (w->has_stepped) {
        // Bring the device states back before the dense arrays grow.
        std::vector<PmhRigidBodyState> cur(w->body_count);
        PmhResult r = pmhWorldCopyStates(world, cur.data());
        if (r != PMH_SUCCESS) return r;
        w->host_states = cur;
    }
    w->host_params.insert(w->host_params.end(), parameters, parameters + count);
    w->host_states.insert(w->host_states.end(), states, states + count);
    w->body_count += count;
    w->bodies_dirty = true;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetWarmstart(PmhWorld world, const float *params, uint32_t policy) {
    if (!world || (params && (policy == 0u || (policy & ~3u) != 0u))) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    w->warm_policy = policy;
    if (!params) {
        w->warm_enabled = false;
        return PMH_SUCCESS;
    }
    for (int k = 0; k < pmh::kMlpParams; ++k)
        if (!std::isfinite(params[k])) return PMH_ERROR_INVALID_ARGUMENT;
    if (hipSetDevice(w->device) != hipSuccess) return PMH_ERROR_DEVICE_LOST;
    PmhResult r = w->mlp_dev.reserve(pmh::kMlpParams);
    if (r != PMH_SUCCESS) return r;
    PMH_CHECK(hipMemcpy(w->mlp_dev.p, params, sizeof(float) * pmh::kMlpParams, hipMemcpyHostToDevice));
    w->warm_enabled = true;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetContactLog(PmhWorld world, uint32_t capacity) {
    if (!world) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (hipSetDevice(w->device) != hipSuccess) return PMH_ERROR_DEVICE_LOST;
    w->log_cap = capacity;
    if (capacity) {
        PmhResult r = w->log_dev.reserve(size_t(capacity) * pmh::kMlpSample);
        if (r != PMH_SUCCESS) return r;
    }
    if (w->solver.counters.p) {
        PMH_CHECK(hipDeviceSynchronize());
        const uint32_t zero = 0;
        PMH_CHECK(hipMemcpy(&w->solver.counters.p->log_count, &zero, sizeof zero, hipMemcpyHostToDevice));
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetContactLog(PmhWorld world, float *samples, uint32_t capacity, uint32_t *count) {
    if (!world || !count || (capacity && !samples)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    *count = 0;
    if (!w->has_stepped || !w->log_cap) return PMH_SUCCESS;
    PMH_CHECK(hipStreamSynchronize(w->stream));
    uint32_t logged = 0;
    PMH_CHECK(hipMemcpy(&logged, &w->solver.counters.p->log_count, sizeof logged, hipMemcpyDeviceToHost));
    logged = std::min(logged, w->log_cap);
    *count = logged;
    const uint32_t copy = std::min(logged, capacity);
    if (copy)
        PMH_CHECK(hipMemcpy(samples, w->log_dev.p, sizeof(float) * size_t(copy) * pmh::kMlpSample, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}

PmhResult pmhWorldWarmstartStats(PmhWorld world, uint64_t *steps, uint64_t *fallbacks) {
    if (!world || !steps || !fallbacks) return PMH_ERROR_INVALID_ARGUMENT;
    *steps = W(world)->warm_steps;
    *fallbacks = W(world)->warm_fallbacks;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetSolveLanes(PmhWorld world, uint32_t lanes) {
    if (!world || (lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)) return PMH_ERROR_INVALID_ARGUMENT;
    W(world)->solve_lanes = lanes;
    return PMH_SUCCESS;
}

PmhResult pmhWorldAddScene(PmhWorld world, uint32_t count, const PmhBodyParameters *parameters,
                           const PmhRigidBodyState *states, uint32_t *scene) {
    if (!world || !scene) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    // The first scene is the one plain pmhWorldAddBodies fills; an empty
    // first scene is reused.
    const bool reuse = w->scene_first.size() == 1 && w->body_count == 0;
    if (!reuse) w->scene_first.push_back(w->body_count);
    PmhResult r = pmhWorldAddBodies(world, count, parameters, states);
    if (r != PMH_SUCCESS) {
        if (!reuse) w->scene_first.pop_back();
        return r;
    }
    *scene = uint32_t(w->scene_first.size() - 1);
    return PMH_SUCCESS;
}

uint32_t pmhWorldSceneCount(PmhWorld world) { return world ? uint32_t(W(world)->scene_first.size()) : 0u; }

uint32_t pmhWorldBodyCount(PmhWorld world) { return world ? W(world)->body_count : 0; }

// Worlds of fewer than kMinStepBodies bodies are refused: the first GPU
// steps of a two- and a three-body world with no active pair did not return
// on an earlier build of the step, and no cause has been shown.
constexpr uint32_t kMinStepBodies = 4;

PmhResult pmhWorldStep(PmhWorld world, uint32_t steps) {
    if (!world) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (w->body_count == 0 || steps == 0) return PMH_SUCCESS;
    if (w->body_count < kMinStepBodies) return PMH_ERROR_FEATURE_NOT_PRESENT;
    PmhResult r = w->ensure_device();
    if (r != PMH_SUCCESS) return r;
    // Substeps are enqueued back to back; each one's solve launch is
    // bracketed by its own event pair, read once the batch is done.  Every
    // substep reads its status back before its solve (with the colour
    // schedule) and returns at the first overflow or colouring failure; the
    // states then stay those before the failing substep.
    const uint64_t total = uint64_t(steps) * w->substeps;
    const uint32_t batch = 256;
    for (uint64_t done = 0; done < total; done += batch) {
        const uint32_t count = uint32_t(std::min<uint64_t>(batch, total - done));
        if (w->timing_enabled && w->solve_events.size() < 2 * size_t(count)) {
            const size_t old = w->solve_events.size();
            w->solve_events.resize(2 * size_t(count), nullptr);
            for (size_t k = old; k < w->solve_events.size(); ++k) PMH_CHECK(hipEventCreate(&w->solve_events[k]));
        }
        const bool fallback = w->warm_enabled && w->residual_tolerance > 0.0f;
        for (uint32_t k = 0; k < count; ++k) {
            r = fallback ? w->step_with_fallback(k) : w->step_once(k);
            if (r != PMH_SUCCESS) {
                w->has_stepped = true;   // earlier substeps of the batch did run
                const PmhResult s = pmh::sync_stream(w->stream);
                return s != PMH_SUCCESS ? s : r;
            }
        }
        r = pmh::sync_stream(w->stream);
        if (r != PMH_SUCCESS) return r;
        if (w->timing_enabled) {
            for (uint32_t k = 0; k < count; ++k) {
                float ms = 0.0f;
                PMH_CHECK(hipEventElapsedTime(&ms, w->solve_events[2 * k], w->solve_events[2 * k + 1]));
                w->solve_timing.total_ms += ms;
            }
            w->last_solve_ms = 0.0f;
            PMH_CHECK(hipEventElapsedTime(&w->last_solve_ms, w->solve_events[2 * (count - 1)],
                                          w->solve_events[2 * (count - 1) + 1]));
        }
        r = w->kernel_timer.collect();
        if (r != PMH_SUCCESS) return r;
    }
    w->has_stepped = true;
    pmh::StepCounters c;
    PMH_CHECK(hipMemcpy(&c, w->solver.counters.p, sizeof c, hipMemcpyDeviceToHost));
    if (c.sweep_failed) return pmh::report_sweep_failure(w->stream, w->solver.counters.p);
    return pmh::status_result(c.status);
}

PmhResult pmhWorldSetGravity(PmhWorld world, const float gravity[3]) {
    if (!world || !gravity || !std::isfinite(gravity[0]) || !std::isfinite(gravity[1]) || !std::isfinite(gravity[2]))
        return PMH_ERROR_INVALID_ARGUMENT;
    W(world)->gravity = {gravity[0], gravity[1], gravity[2]};
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetSweepWaitLimit(PmhWorld world, uint32_t polls) {
    if (!world) return PMH_ERROR_INVALID_ARGUMENT;
    W(world)->sweep_wait_limit = polls;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetColorWaitLimit(PmhWorld world, uint32_t polls) {
    if (!world) return PMH_ERROR_INVALID_ARGUMENT;
    W(world)->color_wait_limit = polls;
    return PMH_SUCCESS;
}

PmhResult pmhWorldApplyRigidInputs(PmhWorld world, uint32_t count, const PmhRigidInput *inputs) {
    if (!world || (count && !inputs)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    const uint32_t n = w->body_count;
    for (uint32_t k = 0; k < count; ++k) {
        const PmhRigidInput &in = inputs[k];
        if (in.body_index >= n || w->host_params[in.body_index].motion != pmh::kDynamic)
            return PMH_ERROR_INVALID_ARGUMENT;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(in.force[c]) || !std::isfinite(in.torque[c]) ||
                !std::isfinite(in.linear_impulse[c]) || !std::isfinite(in.angular_impulse[c]))
                return PMH_ERROR_INVALID_ARGUMENT;
    }
    w->host_inputs.resize(4 * size_t(n), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    w->host_targets.resize(n);
    w->host_target_flags.resize(n, 0u);
    for (uint32_t k = 0; k < count; ++k) {
        const PmhRigidInput &in = inputs[k];
        const float *v[4] = {in.force, in.torque, in.linear_impulse, in.angular_impulse};
        for (int q = 0; q < 4; ++q) {
            float4 &a = w->host_inputs[4 * size_t(in.body_index) + q];
            a.x = a.x + v[q][0];
            a.y = a.y + v[q][1];
            a.z = a.z + v[q][2];
        }
    }
    w->inputs_dirty = true;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetKinematicTargets(PmhWorld world, uint32_t first, uint32_t count,
                                      const PmhRigidBodyState *targets) {
    if (!world || (count && !targets)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    const uint32_t n = w->body_count;
    if (uint64_t(first) + count > n) return PMH_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k) {
        if (w->host_params[first + k].motion != pmh::kKinematic) return PMH_ERROR_INVALID_ARGUMENT;
        const float *f = reinterpret_cast<const float *>(&targets[k]);
        for (size_t c = 0; c < sizeof(PmhRigidBodyState) / sizeof(float); ++c)
            if (!std::isfinite(f[c])) return PMH_ERROR_INVALID_ARGUMENT;
        const PmhQuat o = targets[k].orientation;
        if (o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w <= pmh::kEpsilon * pmh::kEpsilon)
            return PMH_ERROR_INVALID_ARGUMENT;
    }
    w->host_inputs.resize(4 * size_t(n), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    w->host_targets.resize(n);
    w->host_target_flags.resize(n, 0u);
    for (uint32_t k = 0; k < count; ++k) {
        PmhRigidBodyState t = targets[k];
        const pmh::Q q = pmh::qnormalize({t.orientation.x, t.orientation.y, t.orientation.z, t.orientation.w});
        t.orientation = {q.x, q.y, q.z, q.w};
        w->host_targets[first + k] = t;
        w->host_target_flags[first + k] = 1u;
    }
    w->inputs_dirty = true;
    return PMH_SUCCESS;
}

// add_rigid_constraint (runtime_rigid.cu.inc:552-596): the options with
// normalised local frames, a fresh state enabled as the options say, zero
// rows, alive.
static PmhRigidConstraint new_joint(const PmhRigidConstraintOptions &o, uint32_t generation) {
    PmhRigidConstraint j;
    std::memset(&j, 0, sizeof j);
    j.options = o;
    for (PmhQuat *q : {&j.options.local_orientation_a, &j.options.local_orientation_b}) {
        const pmh::Q n = pmh::qnormalize(pmh::Q{q->x, q->y, q->z, q->w});
        *q = {n.x, n.y, n.z, n.w};
    }
    j.generation = generation ? generation : 1u;
    j.alive = 1u;
    j.state_enabled = o.enabled;
    j.dense_a = j.dense_b = pmh::kInvalid;
    j.avbd_next_a = j.avbd_next_b = pmh::kInvalid;
    return j;
}

PmhResult pmhWorldAddJoints(PmhWorld world, uint32_t count, const PmhRigidConstraintOptions *options) {
    if (!world || (count && !options)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    for (uint32_t k = 0; k < count; ++k)
        if (options[k].type > pmh::kJointMotor) return PMH_ERROR_INVALID_ARGUMENT;
    PmhResult r = w->pull_joints();
    if (r != PMH_SUCCESS) return r;
    for (uint32_t k = 0; k < count; ++k) {
        w->host_joints.push_back(new_joint(options[k], 1u));
    }
    w->joints_dirty = true;
    return PMH_SUCCESS;
}

uint32_t pmhWorldJointCount(PmhWorld world) { return world ? uint32_t(W(world)->host_joints.size()) : 0u; }

PmhResult pmhWorldGetJoints(PmhWorld world, PmhRigidConstraint *joints) {
    if (!world || !joints) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (!w->joints_dirty) {
        PmhResult r = w->pull_joints();
        if (r != PMH_SUCCESS) return r;
    }
    std::copy(w->host_joints.begin(), w->host_joints.end(), joints);
    return PMH_SUCCESS;
}

PmhResult pmhWorldCopyStates(PmhWorld world, PmhRigidBodyState *states) {
    if (!world || !states) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (!w->has_stepped || w->bodies_dirty) {
        std::copy(w->host_states.begin(), w->host_states.end(), states);
        return PMH_SUCCESS;
    }
    PMH_CHECK(hipStreamSynchronize(w->stream));
    PMH_CHECK(hipMemcpy(states, w->state_cur.p, sizeof(PmhRigidBodyState) * w->body_count, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetStates(PmhWorld world, const PmhRigidBodyState *states) {
    if (!world || !states) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    w->host_states.assign(states, states + w->body_count);
    if (w->has_stepped && !w->bodies_dirty) {
        PMH_CHECK(hipMemcpy(w->state_cur.p, states, sizeof(PmhRigidBodyState) * w->body_count, hipMemcpyHostToDevice));
        PMH_CHECK(hipMemset(w->cache_n.p, 0, sizeof(uint32_t) * 2));
        PMH_CHECK(hipMemset(w->solver.prev_vel.p, 0, sizeof(float4) * w->body_count));
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldSetRigidStates(PmhWorld world, uint32_t first, uint32_t count, const PmhRigidBodyState *states) {
    if (!world || (count && !states)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    const uint32_t n = w->body_count;
    if (uint64_t(first) + count > n) return PMH_ERROR_INVALID_ARGUMENT;
    std::vector<PmhRigidBodyState> s(states, states + count);
    for (PmhRigidBodyState &x : s) {
        const float *f = reinterpret_cast<const float *>(&x);
        for (size_t c = 0; c < sizeof x / sizeof(float); ++c)
            if (!std::isfinite(f[c])) return PMH_ERROR_INVALID_ARGUMENT;
        const PmhQuat o = x.orientation;
        if (o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w <= pmh::kEpsilon * pmh::kEpsilon)
            return PMH_ERROR_INVALID_ARGUMENT;
        const pmh::Q q = pmh::qnormalize({o.x, o.y, o.z, o.w});
        x.orientation = {q.x, q.y, q.z, q.w};
    }
    PmhResult r = w->joints_dirty ? PMH_SUCCESS : w->pull_joints();
    if (r != PMH_SUCCESS) return r;
    std::copy(s.begin(), s.end(), w->host_states.begin() + first);
    // Pending inputs and targets of these bodies are dropped.
    for (uint32_t i = first; i < first + count; ++i) {
        if (4 * size_t(i) + 3 < w->host_inputs.size())
            for (int q = 0; q < 4; ++q) w->host_inputs[4 * size_t(i) + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i < w->host_target_flags.size()) w->host_target_flags[i] = 0u;
    }
    bool rows = false;
    for (PmhRigidConstraint &j : w->host_joints)
        if (j.alive && ((j.options.body_a.index >= first && j.options.body_a.index < first + count) ||
                        (j.options.body_b.index >= first && j.options.body_b.index < first + count))) {
            std::memset(j.rows, 0, sizeof j.rows);
            rows = true;
        }
    if (rows) w->joints_dirty = true;
    if (w->has_stepped && !w->bodies_dirty && count) {
        PMH_CHECK(hipStreamSynchronize(w->stream));
        PMH_CHECK(hipMemcpy(w->state_cur.p + first, s.data(), sizeof(PmhRigidBodyState) * count,
                            hipMemcpyHostToDevice));
        PMH_CHECK(hipMemset(w->solver.prev_vel.p + first, 0, sizeof(float4) * count));
        PMH_CHECK(hipMemset(w->inputs_dev.p + 4 * size_t(first), 0, sizeof(float4) * 4 * count));
        PMH_CHECK(hipMemset(w->target_flags_dev.p + first, 0, sizeof(uint32_t) * count));
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldWriteJoints(PmhWorld world, uint32_t first, uint32_t count, const PmhRigidConstraintOptions *options,
                              const uint32_t *active) {
    if (!world || (count && (!options || !active))) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (uint64_t(first) + count > w->host_joints.size()) return PMH_ERROR_INVALID_ARGUMENT;
    const uint32_t n = w->body_count;
    for (uint32_t k = 0; k < count; ++k) {
        if (!active[k]) continue;
        const PmhRigidConstraintOptions &o = options[k];
        if (o.type > pmh::kJointMotor || o.body_a.index >= n || o.body_b.index >= n ||
            !std::isfinite(o.breaking_impulse_threshold) || o.breaking_impulse_threshold < 0.0f ||
            o.solver_iterations < 1u || o.solver_iterations > 64u)
            return PMH_ERROR_INVALID_ARGUMENT;
        if (w->host_params[o.body_a.index].motion != pmh::kDynamic &&
            w->host_params[o.body_b.index].motion != pmh::kDynamic)
            return PMH_ERROR_INVALID_ARGUMENT;
    }
    PmhResult r = w->joints_dirty ? PMH_SUCCESS : w->pull_joints();
    if (r != PMH_SUCCESS) return r;
    for (uint32_t k = 0; k < count; ++k) {
        PmhRigidConstraint &j = w->host_joints[first + k];
        if (j.alive) {                                  // remove_rigid_constraint
            j.alive = 0u;
            j.state_enabled = j.state_broken = 0u;
            j.applied_impulse = 0.0f;
            j.generation = j.generation + 1u ? j.generation + 1u : 1u;
        }
        if (active[k]) j = new_joint(options[k], j.generation);
    }
    w->joints_dirty = true;
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetContactEvents(PmhWorld world, PmhContactEvent *events, uint32_t capacity, uint32_t *count) {
    if (!world || !count || (capacity && !events)) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    *count = 0;
    if (!w->has_stepped) return PMH_SUCCESS;
    pmh::Solver &v = w->solver;
    pmh::StepCounters c;
    PMH_CHECK(hipStreamSynchronize(w->stream));
    PMH_CHECK(hipMemcpy(&c, v.counters.p, sizeof c, hipMemcpyDeviceToHost));
    const uint32_t M = c.pairs, P = c.points;
    std::vector<uint32_t> m_count, m_start, flags;
    std::vector<uint2> ab;
    std::vector<float4> np, pt, acc, rep;
    PmhResult r;
    if ((r = fetch(ab, v.ab.p, M)) || (r = fetch(m_count, v.m_count.p, M)) || (r = fetch(m_start, v.m_start.p, M)))
        return r;
    if (capacity && ((r = fetch(np, v.c_np.p, P)) || (r = fetch(pt, v.c_pt.p, P)) || (r = fetch(acc, v.c_acc.p, P)) ||
                     (r = fetch(rep, v.c_rep.p, P)) || (r = fetch(flags, v.c_flags.p, P))))
        return r;
    uint32_t cursor = 0;
    for (uint32_t m = 0; m < M; ++m)
        for (uint32_t k = 0; k < m_count[m]; ++k, ++cursor) {
            if (cursor >= capacity) continue;
            const uint32_t i = m_start[m] + k;
            const bool persistent = ((flags[i] >> 1) & 1u) != 0u;
            const float4 imp = persistent ? acc[i] : rep[i];
            PmhContactEvent &e = events[cursor];
            e.type = 0u;
            e.first = ab[m].x;
            e.second = ab[m].y;
            e.position = {pt[i].x, pt[i].y, pt[i].z};
            e.normal = {np[i].x, np[i].y, np[i].z};
            e.penetration = std::max(0.0f, np[i].w);
            e.normal_impulse = imp.x;
            e.friction_impulse = {imp.y, imp.z, imp.w};
        }
    *count = cursor;
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetStats(PmhWorld world, PmhWorldStats *stats) {
    if (!world || !stats) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    std::memset(stats, 0, sizeof *stats);
    if (!w->has_stepped) return PMH_SUCCESS;
    pmh::StepCounters c;
    PMH_CHECK(hipMemcpy(&c, w->solver.counters.p, sizeof c, hipMemcpyDeviceToHost));
    PmhContactSchedule s;
    PMH_CHECK(hipMemcpy(&s, w->solver.schedule.p, sizeof s, hipMemcpyDeviceToHost));
    stats->pairs = c.pairs;
    stats->manifolds = c.manifolds;
    stats->points = c.points;
    stats->colors = c.colors;
    stats->islands = c.islands;
    stats->impact = c.impact;
    stats->budget = s.budget;
    stats->unsupported_pairs = c.unsupported;
    stats->cache_hits = c.cache_hits;
    stats->phases = s.counts[7] != 0 ? 2 : 1;
    stats->status = c.status;
    stats->pair_capacity = w->solver.pair_cap;
    stats->point_capacity = w->solver.point_cap;
    stats->points_allocated = c.point_alloc;
    stats->triangle_pairs = c.tri_pairs;
    stats->fallback_pairs = c.fallback_pairs;
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetPhaseTimes(PmhWorld world, PmhPhaseTimes *times) {
    if (!world || !times) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    std::memset(times, 0, sizeof *times);
    if (!w->has_stepped) return PMH_SUCCESS;
    if (!w->timing_enabled) return PMH_ERROR_FEATURE_NOT_PRESENT;
    for (int k = 0; k < PMH_PHASE_COUNT - 1; ++k)
        PMH_CHECK(hipEventElapsedTime(&times->ms[k], w->events[k], w->events[k + 1]));
    // Split the prepare+solve interval at the solve kernel's own events.
    const float solve = w->last_solve_ms;
    const float prep_solve = times->ms[4];
    times->ms[4] = prep_solve - solve;
    times->ms[6] = times->ms[5];
    times->ms[5] = solve;
    return PMH_SUCCESS;
}

PmhResult pmhWorldSolveTiming(PmhWorld world, float *milliseconds, uint32_t *launches, int reset) {
    if (!world) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    if (milliseconds) *milliseconds = w->solve_timing.total_ms;
    if (launches) *launches = w->solve_timing.launches;
    if (reset) { w->solve_timing.total_ms = 0.0f; w->solve_timing.launches = 0; }
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetResiduals(PmhWorld world, float *out, float *maximum, uint32_t *executed) {
    if (!world || !out || !executed) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    const float nan = std::nanf("");
    if (!w->has_stepped) {
        std::fill(out, out + 2 * pmh::kResidualSlots, nan);
        if (maximum) std::fill(maximum, maximum + 2 * pmh::kResidualSlots, nan);
        executed[0] = executed[1] = 0u;
        return PMH_SUCCESS;
    }
    std::unique_ptr<pmh::ResidualLog> held(new pmh::ResidualLog);
    pmh::ResidualLog &log = *held;
    PMH_CHECK(hipStreamSynchronize(w->stream));
    PMH_CHECK(hipMemcpy(&log, w->solver.residual.p, sizeof log, hipMemcpyDeviceToHost));
    pmh::StepCounters c;
    PMH_CHECK(hipMemcpy(&c, w->solver.counters.p, sizeof c, hipMemcpyDeviceToHost));
    // A phase that did not run (realtime profile, or no impact) reports 0.
    const bool second = w->profile != 1u && c.impact != 0u;   // legacy profile with an impact
    executed[0] = log.executed[0];
    executed[1] = second ? log.executed[1] : 0u;
    for (uint32_t ph = 0; ph < 2; ++ph)
        for (uint32_t it = 0; it < pmh::kResidualSlots; ++it) {
            const uint32_t k = ph * pmh::kResidualSlots + it;
            const pmh::ResidualTotal t = pmh::residual_fold(log, k);
            float v;
            std::memcpy(&v, &t.bits, sizeof v);
            out[k] = it < executed[ph] ? pmh::residual_rms(t.sum, t.count) : nan;
            if (maximum) maximum[k] = it < executed[ph] ? v : nan;
        }
    return PMH_SUCCESS;
}

PmhResult pmhWorldKernelTiming(PmhWorld world, PmhKernelTiming *timing, int reset) {
    if (!world || !timing) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    timing->milliseconds = w->kernel_timer.ms;
    timing->launches = w->kernel_timer.launches;
    timing->algorithmic_bytes = w->kernel_timer.bytes;
    if (reset) {
        w->kernel_timer.ms = 0.0;
        w->kernel_timer.bytes = 0.0;
        w->kernel_timer.launches = 0;
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetTotals(PmhWorld world, PmhWorldTotals *totals, int reset) {
    if (!world || !totals) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    std::memset(totals, 0, sizeof *totals);
    if (!w->solver.totals.p) return PMH_SUCCESS;
    PMH_CHECK(hipStreamSynchronize(w->stream));
    pmh::SolveTotals t;
    PMH_CHECK(hipMemcpy(&t, w->solver.totals.p, sizeof t, hipMemcpyDeviceToHost));
    totals->steps = t.steps;
    totals->body_iterations = t.body_iterations;
    totals->velocity_phases = t.velocity_phases;
    totals->algorithmic_bytes = t.algorithmic_bytes;
    if (reset) PMH_CHECK(hipMemset(w->solver.totals.p, 0, sizeof t));
    return PMH_SUCCESS;
}

uint32_t pmhWorldPairCount(PmhWorld world) {
    if (!world || !W(world)->has_stepped) return 0;
    pmh::StepCounters c;
    if (hipMemcpy(&c, W(world)->solver.counters.p, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return c.pairs;
}

PmhResult pmhWorldGetPairs(PmhWorld world, uint32_t *codes) {
    if (!world || !codes) return PMH_ERROR_INVALID_ARGUMENT;
    const uint32_t m = pmhWorldPairCount(world);
    // Codes a*n+b are the reference's 32-bit pair codes: defined while
    // n*n fits 32 bits.
    const uint64_t nb = W(world)->body_count;
    if (nb * nb > 0xffffffffull) return PMH_ERROR_CAPACITY_EXCEEDED;
    if (m) PMH_CHECK(hipMemcpy(codes, W(world)->solver.code.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetManifolds(PmhWorld world, PmhContactManifold *out) {
    if (!world || !out) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    pmh::Solver &v = w->solver;
    pmh::StepCounters c;
    PMH_CHECK(hipMemcpy(&c, v.counters.p, sizeof c, hipMemcpyDeviceToHost));
    const uint32_t M = c.pairs, P = c.points;
    std::vector<uint32_t> count, start, slot, face;
    std::vector<float4> init, np, pt, aa, ab, err, lam, pen, acc, rep;
    std::vector<uint32_t> flags;
    std::vector<uint2> pair_ab;
    PmhResult r;
    if ((r = fetch(pair_ab, v.ab.p, M)) ||
        (r = fetch(count, v.m_count.p, M)) || (r = fetch(start, v.m_start.p, M)) || (r = fetch(slot, v.m_slot.p, M)) ||
        (r = fetch(face, v.m_face.p, M)) || (r = fetch(init, v.m_init_rel.p, M)) || (r = fetch(np, v.c_np.p, P)) ||
        (r = fetch(pt, v.c_pt.p, P)) || (r = fetch(aa, v.c_aa.p, P)) || (r = fetch(ab, v.c_ab.p, P)) ||
        (r = fetch(err, v.c_err.p, P)) || (r = fetch(lam, v.c_lam.p, P)) || (r = fetch(pen, v.c_pen.p, P)) ||
        (r = fetch(acc, v.c_acc.p, P)) || (r = fetch(rep, v.c_rep.p, P)) || (r = fetch(flags, v.c_flags.p, P)))
        return r;
    // The bodies' hinge contact frames of the step (worlds with joints);
    // without joints a frame carries only the static flag.
    std::vector<PmhHingeContactFrame> frames;
    if (v.nj && w->has_stepped && w->hinge_frames.p && (r = fetch(frames, w->hinge_frames.p, w->body_count))) return r;
    for (uint32_t m = 0; m < M; ++m) {
        PmhContactManifold &o = out[m];
        std::memset(&o, 0, sizeof o);
        o.count = count[m];
        o.initial_relative_position = {init[m].x, init[m].y, init[m].z};
        o.face_patch = face[m];
        o.response_slot = count[m] ? slot[m] : 0xffffffffu;
        for (uint32_t k = 0; k < count[m] && k < 8; ++k) {
            const uint32_t i = start[m] + k;
            PmhContact &t = o.contacts[k];
            t.normal = {np[i].x, np[i].y, np[i].z};
            t.penetration = np[i].w;
            t.point = {pt[i].x, pt[i].y, pt[i].z};
            t.initial_normal_speed = pt[i].w;
            t.avbd_anchor_a = {aa[i].x, aa[i].y, aa[i].z};
            t.impact_fraction = aa[i].w;
            t.avbd_anchor_b = {ab[i].x, ab[i].y, ab[i].z};
            t.avbd_error = {err[i].x, err[i].y, err[i].z};
            t.avbd_dual[0] = {lam[i].x, pen[i].x};
            t.avbd_dual[1] = {lam[i].y, pen[i].y};
            t.avbd_dual[2] = {lam[i].z, pen[i].z};
            t.accumulated_normal_impulse = acc[i].x;
            t.accumulated_friction_impulse = {acc[i].y, acc[i].z, acc[i].w};
            t.reported_normal_impulse = rep[i].x;
            t.reported_friction_impulse = {rep[i].y, rep[i].z, rep[i].w};
            t.avbd_cached = flags[i] & 1u;
            t.persistent = (flags[i] >> 1) & 1u;
            if (!frames.empty()) {
                t.body_hinge = frames[pair_ab[m].x];
                t.collider_hinge = frames[pair_ab[m].y];
            } else {
                t.body_hinge.static_body = w->host_params[pair_ab[m].x].motion == 0u ? 1u : 0u;
                t.collider_hinge.static_body = w->host_params[pair_ab[m].y].motion == 0u ? 1u : 0u;
            }
        }
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetAvbdBodies(PmhWorld world, PmhAvbdBody *out) {
    if (!world || !out) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    pmh::Solver &v = w->solver;
    const uint32_t n = w->body_count;
    std::vector<float4> tpos, tq, tv, tw, disp, dq, inertia, pvel;
    std::vector<uint32_t> color, parent, impact, joint_head;
    PmhResult r;
    if (v.nj && w->has_stepped && (r = fetch(joint_head, v.joint_head.p, n))) return r;
    if ((r = fetch(tpos, v.tpos.p, n)) || (r = fetch(tq, v.tq.p, n)) || (r = fetch(tv, v.tv.p, n)) ||
        (r = fetch(tw, v.tw.p, n)) || (r = fetch(disp, v.disp.p, 2 * size_t(n))) || (r = fetch(dq, v.dq.p, n)) ||
        (r = fetch(inertia, v.inertia.p, 2 * size_t(n))) || (r = fetch(pvel, v.prev_vel.p, n)) ||
        (r = fetch(color, v.color.p, n)) || (r = fetch(parent, v.parent.p, n)) || (r = fetch(impact, v.impact.p, n)))
        return r;
    for (uint32_t i = 0; i < n; ++i) {
        PmhAvbdBody &b = out[i];
        std::memset(&b, 0, sizeof b);
        b.target.position = {tpos[i].x, tpos[i].y, tpos[i].z};
        b.target.orientation = {tq[i].x, tq[i].y, tq[i].z, tq[i].w};
        b.target.linear_velocity = {tv[i].x, tv[i].y, tv[i].z};
        b.target.angular_velocity = {tw[i].x, tw[i].y, tw[i].z};
        b.orientation_delta = {dq[i].x, dq[i].y, dq[i].z, dq[i].w};
        const float d[6] = {disp[2 * i].x, disp[2 * i].y, disp[2 * i].z, disp[2 * i + 1].x, disp[2 * i + 1].y,
                            disp[2 * i + 1].z};
        std::memcpy(b.displacement, d, sizeof d);
        const float4 i0 = inertia[2 * i], i1 = inertia[2 * i + 1];
        b.inertia_h[0] = b.inertia_h[2] = b.inertia_h[5] = i0.x;
        b.inertia_h[9] = i0.y; b.inertia_h[13] = i0.z; b.inertia_h[14] = i0.w;
        b.inertia_h[18] = i1.x; b.inertia_h[19] = i1.y; b.inertia_h[20] = i1.z;
        b.contact_head = b.joint_head = 0xffffffffu;
        if (v.nj && w->has_stepped) b.joint_head = joint_head[i];
        b.color = color[i];
        b.parent = parent[i];
        b.impact = impact[i];
        b.previous_velocity = {pvel[i].x, pvel[i].y, pvel[i].z};
        b.history_valid = pvel[i].w != 0.0f;
    }
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetSchedule(PmhWorld world, PmhContactSchedule *schedule) {
    if (!world || !schedule) return PMH_ERROR_INVALID_ARGUMENT;
    PMH_CHECK(hipMemcpy(schedule, W(world)->solver.schedule.p, sizeof *schedule, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}

PmhResult pmhWorldGetProxies(PmhWorld world, float *bounds) {
    if (!world || !bounds) return PMH_ERROR_INVALID_ARGUMENT;
    World *w = W(world);
    const uint32_t n = w->body_count;
    std::vector<float4> mn, mx;
    PmhResult r;
    if ((r = fetch(mn, w->proxy_mn.p, n)) || (r = fetch(mx, w->proxy_mx.p, n))) return r;
    for (uint32_t i = 0; i < n; ++i) {
        const float v[6] = {mn[i].x, mn[i].y, mn[i].z, mx[i].x, mx[i].y, mx[i].z};
        std::memcpy(bounds + 6 * i, v, sizeof v);
    }
    return PMH_SUCCESS;
}

PmhResult pmhDeviceAlloc(uint64_t bytes, void **pointer) {
    if (!pointer) return PMH_ERROR_INVALID_ARGUMENT;
    PMH_CHECK(hipMalloc(pointer, bytes ? bytes : 1));
    return PMH_SUCCESS;
}
PmhResult pmhDeviceFree(void *pointer) {
    PMH_CHECK(hipFree(pointer));
    return PMH_SUCCESS;
}
PmhResult pmhCopyToDevice(void *device, const void *host, uint64_t bytes) {
    PMH_CHECK(hipMemcpy(device, host, bytes, hipMemcpyHostToDevice));
    return PMH_SUCCESS;
}
PmhResult pmhCopyToHost(void *host, const void *device, uint64_t bytes) {
    PMH_CHECK(hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost));
    return PMH_SUCCESS;
}
PmhResult pmhDeviceSynchronize(void) {
    PMH_CHECK(hipDeviceSynchronize());
    return PMH_SUCCESS;
}
PmhResult pmhDeviceInfo(char *name, uint32_t name_bytes, uint32_t *compute_units) {
    int dev = 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PMH_ERROR_FEATURE_NOT_PRESENT;
    PMH_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PMH_CHECK(hipGetDeviceProperties(&prop, dev));
    if (name && name_bytes) {
        std::strncpy(name, prop.gcnArchName, name_bytes - 1);
        name[name_bytes - 1] = 0;
    }
    if (compute_units) *compute_units = uint32_t(prop.multiProcessorCount);
    return PMH_SUCCESS;
}

}  // extern "C"
