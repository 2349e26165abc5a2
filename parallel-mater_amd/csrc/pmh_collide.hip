// SPDX-License-Identifier: MIT
// Predictor, proxies, broadphase, narrowphase and the contact cache on
// gfx950, and the world step that chains them with the AVBD solve.
//
// Broadphase: the reference filters every proxy overlap (inclusive float
// AABB test) with the rules of pm_filter_rigid_spatial_candidates and
// orders pairs by (body, identity) (geometry.slang:829-947).  Here the
// proxies of bodies no larger than the largest dynamic proxy go into a
// uniform hash grid (cell = 1.0625 x that extent, keyed by the cell of the
// proxy's minimum corner, so any overlapping pair is found among the 27
// neighbouring cells); larger static proxies (floors, walls) form a short
// list every query scans.  Each dynamic body sorts its kept candidates, and
// an exclusive scan over the per-body counts emits the pairs in ascending
// code a*n+b.  The overlap test is the same comparison on the same floats,
// so the pair set and order equal the reference's bit for bit.
//
// Narrowphase: one thread per pair; convex meshes with <= 96 indices take
// the reference's SAT + face clipping path (rigid_convex_face_manifold and
// its swept variant, geometry.slang:1606-1853); analytic spheres/capsules
// use closed forms.  Pairs the reference would send to its triangle-pair
// path are counted and left without contacts.
#include "pmh_internal.hpp"
#include "pmh_world.hpp"

#include <algorithm>

namespace pmh {

__device__ __forceinline__ V3 sv(const PmhVec3 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ Q sq(const PmhQuat &q) { return {q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ PmhVec3 pv(V3 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ PmhQuat pq(Q q) { return {q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ V3 v3(float4 f) { return {f.x, f.y, f.z}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 sarr(const float *a) { return {a[0], a[1], a[2]}; }
__device__ __forceinline__ V3 transform_point(const PmhRigidBodyState &s, V3 p) {
    return add(sv(s.position), rotate(sq(s.orientation), p));
}

// ---------------------------------------------------------------------
// Predictor + proxies (geometry.slang:524-612, 716-827).
// ---------------------------------------------------------------------
__device__ __forceinline__ float correction_reach(const PmhBodyParameters &p, const ShapeDev &s) {
    if (p.motion != kDynamic) return 0.0f;
    const V3 e = sub(sarr(s.mx), sarr(s.mn));
    return smin(0.5f, 0.25f * smax(0.0f, smin(e.x, smin(e.y, e.z))));
}

__device__ __forceinline__ V3 inverse_inertia_world(const PmhBodyParameters &p, Q o, V3 w) {
    const V3 l = inv_rotate(o, w);
    return rotate(o, {l.x * p.inverse_inertia_local.x, l.y * p.inverse_inertia_local.y,
                      l.z * p.inverse_inertia_local.z});
}

__global__ void k_integrate_proxy(uint32_t n, const PmhBodyParameters *P, const ShapeDev *shapes,
                                  const PmhRigidBodyState *cur, PmhRigidBodyState *prev,
                                  PmhRigidBodyState *pred, float4 *pmn, float4 *pmx, V3 g, float dt,
                                  StepCounters *ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PmhBodyParameters b = P[i];
    const PmhRigidBodyState s0 = cur[i];
    prev[i] = s0;
    PmhRigidBodyState s = s0;
    if (b.motion == kStatic || b.motion == kKinematic) {
        s.linear_velocity = {0.0f, 0.0f, 0.0f};
        s.angular_velocity = {0.0f, 0.0f, 0.0f};
    } else {
        const V3 zero = {0.0f, 0.0f, 0.0f};
        V3 v = add(sv(s.linear_velocity), mul(add(g, mul(zero, b.inverse_mass)), dt));
        V3 w = add(sv(s.angular_velocity), mul(inverse_inertia_world(b, sq(s.orientation), zero), dt));
        v = add(v, mul(zero, b.inverse_mass));
        w = add(w, inverse_inertia_world(b, sq(s.orientation), zero));
        v = mul(v, 1.0f / (1.0f + b.linear_damping * dt));
        w = mul(w, 1.0f / (1.0f + b.angular_damping * dt));
        v = clamp_length(v, b.maximum_linear_speed);
        w = clamp_length(w, b.maximum_angular_speed);
        const V3 p = add(sv(s.position), mul(v, dt));
        const Q o = sq(s.orientation);
        const Q d = qmul({w.x, w.y, w.z, 0.0f}, o);
        const Q no = qnormalize({o.x + 0.5f * d.x * dt, o.y + 0.5f * d.y * dt, o.z + 0.5f * d.z * dt,
                                 o.w + 0.5f * d.w * dt});
        s.position = pv(p);
        s.orientation = pq(no);
        s.linear_velocity = pv(v);
        s.angular_velocity = pv(w);
    }
    pred[i] = s;
    // Proxy: production arguments path_deviation 0, correction_budget 1,
    // epsilon 1e-5 (production_realtime_backend.cu.inc:2663-2684).
    const ShapeDev sh = shapes[b.mesh_index];
    const float eps = 1.0e-5f;
    const float reach = len(sarr(sh.center)) + sh.bradius + b.collision_margin + eps;
    const float cr = correction_reach(b, sh);
    const V3 bc = {smin(1.0f, cr), smin(1.0f, cr), smin(1.0f, cr)};
    const V3 e = add(add(V3{0.0f, 0.0f, 0.0f}, bc), V3{reach, reach, reach});
    const V3 mn = sub(cmin(sv(s0.position), sv(s.position)), e);
    const V3 mx = add(cmax(sv(s0.position), sv(s.position)), e);
    pmn[i] = f4(mn, 0.0f);
    pmx[i] = f4(mx, 0.0f);
    (void)ctr;
}

// ---------------------------------------------------------------------
// Broadphase.
// ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_hash(int x, int y, int z) {
    uint32_t h = uint32_t(x) * 0x8da6b343u ^ uint32_t(y) * 0xd8163841u ^ uint32_t(z) * 0xcb1ab31fu;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return h;
}

constexpr int kLargeCell = 0x7fffffff;

// Proxies no wider than `cell` go into the hash grid under the cell of
// their minimum corner; wider ones (floors, walls, fast movers) into the
// large list.  Any overlapping gridded pair then lies in neighbouring cells.
__global__ void k_bp_cells(uint32_t n, const float4 *pmn, const float4 *pmx, int4 *cell_of, uint32_t *bucket_count,
                           uint32_t table_mask, uint32_t *large, StepCounters *ctr, float cell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 mn = pmn[i], mx = pmx[i];
    const float ext = smax(mx.x - mn.x, smax(mx.y - mn.y, mx.z - mn.z));
    if (!(ext <= cell) || !isfinite(mn.x) || !isfinite(mn.y) || !isfinite(mn.z)) {
        large[atomicAdd(&ctr->large_count, 1u)] = i;
        cell_of[i] = make_int4(kLargeCell, kLargeCell, kLargeCell, 0);
        return;
    }
    const int cx = int(floorf(mn.x / cell)), cy = int(floorf(mn.y / cell)), cz = int(floorf(mn.z / cell));
    const uint32_t b = cell_hash(cx, cy, cz) & table_mask;
    cell_of[i] = make_int4(cx, cy, cz, int(b));
    atomicAdd(&bucket_count[b], 1u);
}

__global__ void k_bp_scatter(uint32_t n, const int4 *cell_of, const uint32_t *bucket_start, uint32_t *cursor,
                             uint32_t *entries) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 c = cell_of[i];
    if (c.x == kLargeCell) return;
    entries[bucket_start[c.w] + atomicAdd(&cursor[c.w], 1u)] = i;
}

__device__ __forceinline__ bool overlap(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z &&
           amx.z >= bmn.z;
}

// Kept candidates of dynamic body a.  Pass 0 counts them; pass 1 writes
// them at the body's offset (from an exclusive scan of the counts) and
// sorts the segment ascending (identity order = index), so a body may own
// any number of pairs.
template <int kPass>
__global__ void k_bp_query(uint32_t n, const PmhBodyParameters *P, const float4 *pmn, const float4 *pmx,
                           const int4 *cell_of, const uint32_t *bucket_start, const uint32_t *entries,
                           uint32_t table_mask, const uint32_t *large, uint32_t *cand_count,
                           const uint32_t *offset, uint32_t *code, uint2 *ab, const StepCounters *ctr) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    if (P[a].motion != kDynamic) {
        if (kPass == 0) cand_count[a] = 0;
        return;
    }
    if (kPass == 1 && ctr->pairs == 0) return;   // pair pool overflow: nothing is emitted
    const float4 amn = pmn[a], amx = pmx[a];
    const uint32_t base = kPass == 1 ? offset[a] : 0;
    uint32_t count = 0;
    auto consider = [&](uint32_t b) {
        if (b == a) return;
        if (P[b].motion == kDynamic && !(a < b)) return;
        if (!overlap(amn, amx, pmn[b], pmx[b])) return;
        if (kPass == 1) code[base + count] = b;
        ++count;
    };
    const int4 c = cell_of[a];
    if (c.x == kLargeCell) return;   // k_bp_large owns this body
    {
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = c.x + dx, y = c.y + dy, z = c.z + dz;
                    const uint32_t bk = cell_hash(x, y, z) & table_mask;
                    for (uint32_t e = bucket_start[bk]; e < bucket_start[bk + 1]; ++e) {
                        const uint32_t b = entries[e];
                        const int4 cb = cell_of[b];
                        if (cb.x != x || cb.y != y || cb.z != z) continue;
                        consider(b);
                    }
                }
    }
    const uint32_t nl = ctr->large_count;
    for (uint32_t k = 0; k < nl; ++k) consider(large[k]);
    if (kPass == 0) {
        cand_count[a] = count;
        return;
    }
    uint32_t *out = code + base;
    for (uint32_t k = 1; k < count; ++k) {
        const uint32_t v = out[k];
        uint32_t j = k;
        while (j > 0 && out[j - 1] > v) { out[j] = out[j - 1]; --j; }
        out[j] = v;
    }
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t b = out[k];
        ab[base + k] = make_uint2(a, b);
        out[k] = a * n + b;
    }
}

__global__ void k_set_pairs(StepCounters *ctr, const uint32_t *offset, uint32_t n, uint32_t pair_cap) {
    if (threadIdx.x != 0) return;
    const uint32_t M = offset[n];
    if (M > pair_cap) {
        atomicOr(&ctr->status, kStatusPairOverflow);
        ctr->pairs = 0;
    } else {
        ctr->pairs = M;
    }
}

// Large bodies: one block scans every body in index order; a block-wide
// prefix sum places the hits, so the segment comes out ascending.
template <int kPass>
__global__ void __launch_bounds__(256) k_bp_large(uint32_t n, const PmhBodyParameters *P, const float4 *pmn,
                                                  const float4 *pmx, const uint32_t *large, uint32_t *cand_count,
                                                  const uint32_t *offset, uint32_t *code, uint2 *ab,
                                                  const StepCounters *ctr) {
    __shared__ uint32_t warp_tot[4];
    const uint32_t nl = ctr->large_count;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t k = blockIdx.x; k < nl; k += gridDim.x) {
        const uint32_t a = large[k];
        if (P[a].motion != kDynamic) {
            if (kPass == 0 && threadIdx.x == 0) cand_count[a] = 0;
            continue;
        }
        if (kPass == 1 && ctr->pairs == 0) continue;
        const float4 amn = pmn[a], amx = pmx[a];
        uint32_t running = kPass == 1 ? offset[a] : 0u;
        for (uint32_t base = 0; base < n; base += 256) {
            const uint32_t b = base + threadIdx.x;
            bool hit = false;
            if (b < n && b != a && !(P[b].motion == kDynamic && !(a < b)))
                hit = overlap(amn, amx, pmn[b], pmx[b]);
            const uint64_t bal = __ballot(hit);
            if (lane == 0) warp_tot[wid] = __popcll(bal);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < 4; ++w) {
                if (w < wid) before += warp_tot[w];
                total += warp_tot[w];
            }
            if (kPass == 1 && hit) {
                const uint32_t pos = running + before + __popcll(bal & ((1ull << lane) - 1ull));
                code[pos] = a * n + b;
                ab[pos] = make_uint2(a, b);
            }
            running += total;
            __syncthreads();
        }
        if (kPass == 0 && threadIdx.x == 0) cand_count[a] = running;
    }
}

// ---------------------------------------------------------------------
// Narrowphase.
// ---------------------------------------------------------------------
struct Contact {
    V3 normal, point;
    float pen, impact_fraction;
};
struct Manifold {
    Contact c[8];
    uint32_t count;
    uint32_t face_patch;
};

struct ShapeView {
    ShapeDev s;
    const V3 *v;
    const uint32_t *idx;
    const float4 *planes;
    __device__ float4 plane(uint32_t f) const { return planes[s.plane_off + f]; }
    __device__ V3 vert(uint32_t k) const { return v[s.vtx_off + k]; }
    __device__ uint32_t index(uint32_t k) const { return idx[s.idx_off + k]; }
};

// rigid_add_contact, geometry.slang:1129-1220 (unguided, no swept-witness
// canonicalisation on the convex face path).
__device__ void add_contact(Manifold &m, const Contact &c, float spacing) {
    const float sp2 = spacing * spacing;
    for (uint32_t i = 0; i < m.count; ++i) {
        if (len2(sub(m.c[i].point, c.point)) < sp2) {
            if (c.pen - m.c[i].pen > 0.0f) m.c[i] = c;
            return;
        }
    }
    if (m.count < 8u) { m.c[m.count++] = c; return; }
    uint32_t sh = 0;
    for (uint32_t i = 1; i < m.count; ++i) if (m.c[i].pen < m.c[sh].pen) sh = i;
    if (c.pen > m.c[sh].pen) m.c[sh] = c;
}

// rigid_reduce_face_patch, geometry.slang:1228-1270.
__device__ void reduce_face_patch(Manifold &m, float tol) {
    const float t2 = tol * tol;
    bool removed = true;
    while (removed) {
        removed = false;
        for (uint32_t cand = 0; cand < m.count && !removed; ++cand) {
            const V3 p = m.c[cand].point;
            for (uint32_t f = 0; f < m.count && !removed; ++f) {
                if (f == cand) continue;
                for (uint32_t s = f + 1; s < m.count; ++s) {
                    if (s == cand) continue;
                    const V3 e = sub(m.c[s].point, m.c[f].point);
                    const float e2 = len2(e);
                    if (e2 <= t2) continue;
                    const float t = dot(sub(p, m.c[f].point), e) / e2;
                    if (t <= 1.0e-4f || t >= 0.9999f) continue;
                    const V3 cl = add(m.c[f].point, mul(e, t));
                    if (len2(sub(p, cl)) > t2) continue;
                    for (uint32_t k = cand + 1; k < m.count; ++k) m.c[k - 1] = m.c[k];
                    --m.count;
                    removed = true;
                    break;
                }
            }
        }
    }
}

__device__ float rotational_motion_bound(const PmhRigidBodyState &p, const PmhRigidBodyState &c,
                                         const ShapeDev &s) {
    const Q d = relative_quaternion(sq(p.orientation), sq(c.orientation), false);
    const float sh = len({d.x, d.y, d.z});
    const V3 ma = {smax(fabsf(s.mn[0]), fabsf(s.mx[0])), smax(fabsf(s.mn[1]), fabsf(s.mx[1])),
                   smax(fabsf(s.mn[2]), fabsf(s.mx[2]))};
    return 2.0f * len(ma) * sh;
}

// rigid_convex_face_manifold, geometry.slang:1606-1745.
__device__ bool convex_face_manifold(const PmhRigidBodyState &bs, const ShapeView &bm,
                                     const PmhRigidBodyState &cs, const ShapeView &cm, float margin,
                                     float reach_corr, Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const float rest = smin(margin, 1.0e-3f);
    const float reach = margin + reach_corr;
    float best = -3.402823466e+38f;
    uint32_t ref_face = 0;
    bool ref_is_body = false;
    for (uint32_t side = 0; side < 2; ++side) {
        const ShapeView &rm = side == 0 ? cm : bm;
        const PmhRigidBodyState &rs = side == 0 ? cs : bs;
        const ShapeView &im = side == 0 ? bm : cm;
        const PmhRigidBodyState &is = side == 0 ? bs : cs;
        const Q rq = sq(rs.orientation), iq = sq(is.orientation);
        const V3 rel = sub(sv(is.position), sv(rs.position));
        for (uint32_t f = 0; f < rm.s.idx_count / 3; ++f) {
            const float4 pl = rm.plane(f);
            const V3 nrm = rotate(rq, v3(pl));
            const V3 inn = inv_rotate(iq, nrm);
            float support = 3.402823466e+38f;
            for (uint32_t k = 0; k < im.s.vtx_count; ++k) support = smin(support, dot(inn, im.vert(k)));
            const float sep = support + dot(nrm, rel) - pl.w;
            if (sep > best) { best = sep; ref_face = f; ref_is_body = side != 0; }
        }
    }
    if (best > reach) return true;
    const ShapeView &rm = ref_is_body ? bm : cm;
    const PmhRigidBodyState &rs = ref_is_body ? bs : cs;
    const ShapeView &im = ref_is_body ? cm : bm;
    const PmhRigidBodyState &is = ref_is_body ? cs : bs;
    const Q rq = sq(rs.orientation);
    const float4 reference = rm.plane(ref_face);
    const V3 rn = v3(reference);
    const V3 outward = rotate(rq, rn);
    const V3 iaxis = inv_rotate(sq(is.orientation), outward);
    float alignment = 1.0f;
    uint32_t inc_face = 0;
    for (uint32_t f = 0; f < im.s.idx_count / 3; ++f) {
        const float val = dot(iaxis, v3(im.plane(f)));
        if (val < alignment) { alignment = val; inc_face = f; }
    }
    const V3 normal = mul(outward, ref_is_body ? -1.0f : 1.0f);
    const float4 incident = im.plane(inc_face);
    const uint32_t rfaces = rm.s.idx_count / 3;
    for (uint32_t t = 0; t < im.s.idx_count / 3; ++t) {
        const float4 face = im.plane(t);
        if (dot(v3(face), v3(incident)) < 0.99999f || fabsf(face.w - incident.w) > 1.0e-5f) continue;
        V3 poly[36], clipped[36];
        uint32_t count = 3;
        for (uint32_t c = 0; c < 3; ++c)
            poly[c] = inv_rotate(rq, sub(transform_point(is, im.vert(im.index(t * 3 + c))), sv(rs.position)));
        for (uint32_t pi = 0; pi < rfaces && count != 0; ++pi) {
            const float4 pl = rm.plane(pi);
            const V3 pn = v3(pl);
            const float off = pl.w + (dot(pn, rn) > 0.99999f ? reach : 0.0f);
            uint32_t cc = 0;
            V3 prev = poly[count - 1];
            float pd = dot(pn, prev) - off;
            for (uint32_t k = 0; k < count; ++k) {
                const V3 cur = poly[k];
                const float d = dot(pn, cur) - off;
                if ((d <= 0.0f) != (pd <= 0.0f) && cc < 36)
                    clipped[cc++] = add(prev, mul(sub(cur, prev), pd / (pd - d)));
                if (d <= 0.0f && cc < 36) clipped[cc++] = cur;
                prev = cur;
                pd = d;
            }
            count = cc;
            for (uint32_t k = 0; k < count; ++k) poly[k] = clipped[k];
        }
        for (uint32_t k = 0; k < count; ++k) {
            const float d = dot(rn, poly[k]) - reference.w;
            if (d > reach) continue;
            Contact c;
            c.normal = normal;
            c.point = transform_point(rs, sub(poly[k], mul(rn, d * 0.5f)));
            c.pen = rest - d + 1.0e-5f;
            c.impact_fraction = 1.0f;
            add_contact(out, c, smax(margin * 2.0f, 1.0e-4f));
        }
    }
    if (out.count == 0u) return false;
    reduce_face_patch(out, smax(1.0e-6f, margin * 1.0e-3f));
    out.face_patch = 1u;
    return true;
}

// rigid_swept_convex_face_manifold, geometry.slang:1751-1853.
__device__ bool swept_convex_face_manifold(const PmhRigidBodyState &pb, const PmhRigidBodyState &b,
                                           const ShapeView &bm, const PmhRigidBodyState &pc,
                                           const PmhRigidBodyState &c, const ShapeView &cm, float margin,
                                           Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const Q bd = relative_quaternion(sq(pb.orientation), sq(b.orientation), false);
    const Q cd = relative_quaternion(sq(pc.orientation), sq(c.orientation), false);
    if (bd.x != 0.0f || bd.y != 0.0f || bd.z != 0.0f || cd.x != 0.0f || cd.y != 0.0f || cd.z != 0.0f)
        return false;
    const V3 bmo = sub(sv(b.position), sv(pb.position)), cmo = sub(sv(c.position), sv(pc.position));
    const float rest = smin(margin, 1.0e-3f);
    float entry = 0.0f, exit_time = 1.0f;
    V3 entry_normal = {0.0f, 0.0f, 0.0f};
    for (uint32_t side = 0; side < 2; ++side) {
        const ShapeView &rm = side == 0 ? cm : bm;
        const ShapeView &im = side == 0 ? bm : cm;
        const PmhRigidBodyState &rs = side == 0 ? pc : pb;
        const PmhRigidBodyState &is = side == 0 ? pb : pc;
        const V3 rel = side == 0 ? sub(bmo, cmo) : sub(cmo, bmo);
        for (uint32_t f = 0; f < rm.s.idx_count / 3; ++f) {
            const float4 pl = rm.plane(f);
            const V3 nrm = rotate(sq(rs.orientation), v3(pl));
            const V3 inn = inv_rotate(sq(is.orientation), nrm);
            float support = 3.402823466e+38f;
            for (uint32_t k = 0; k < im.s.vtx_count; ++k) support = smin(support, dot(inn, im.vert(k)));
            const float gap = support + dot(nrm, sub(sv(is.position), sv(rs.position))) - pl.w;
            const float speed = dot(nrm, rel);
            if (speed == 0.0f) {
                if (gap > rest) return true;
            } else {
                const float crossing = (rest - gap) / speed;
                if (speed < 0.0f && crossing > entry) {
                    entry = crossing;
                    entry_normal = mul(nrm, side == 0 ? 1.0f : -1.0f);
                }
                if (speed > 0.0f) exit_time = smin(exit_time, crossing);
            }
        }
    }
    if (entry > exit_time || entry > 1.0f || exit_time < 0.0f) return true;
    if (entry == 0.0f) {
        const bool clipped = convex_face_manifold(b, bm, c, cm, margin, 0.0f, out);
        if (clipped && out.count != 0u) return true;
        if (!convex_face_manifold(pb, bm, pc, cm, margin, 0.0f, out) || out.count == 0u) return false;
        for (uint32_t p = 0; p < out.count; ++p) {
            const float closing = -dot(sub(bmo, cmo), out.c[p].normal);
            if (closing <= 0.0f) return false;
            out.c[p].pen += closing;
            out.c[p].impact_fraction = 0.0f;
        }
        return true;
    }
    PmhRigidBodyState ib = pb, ic = pc;
    ib.position = pv(add(sv(pb.position), mul(bmo, entry)));
    ic.position = pv(add(sv(pc.position), mul(cmo, entry)));
    if (!convex_face_manifold(ib, bm, ic, cm, rest + 1.0e-5f, 0.0f, out) || out.count == 0u) return false;
    for (uint32_t p = 0; p < out.count; ++p) {
        Contact &ct = out.c[p];
        if (dot(ct.normal, entry_normal) < 0.99999f) return false;
        const float remaining = -dot(sub(bmo, cmo), ct.normal) * (1.0f - entry);
        ct.pen = smax(0.0f, remaining + ct.pen) + rest + 1.0e-5f;
        ct.impact_fraction = entry;
    }
    return true;
}

__device__ bool requires_swept(const PmhRigidBodyState &pa, const PmhRigidBodyState &a, const ShapeDev &ma,
                               const PmhRigidBodyState &pb, const PmhRigidBodyState &b, const ShapeDev &mb,
                               float threshold) {
    const bool planes_a = ma.kind != kShapeMesh || ma.has_planes;
    const bool planes_b = mb.kind != kShapeMesh || mb.has_planes;
    if (!planes_a || !planes_b) threshold = smin(threshold, 1.0e-3f);
    const V3 rel = sub(sub(sv(a.position), sv(pa.position)), sub(sv(b.position), sv(pb.position)));
    return len(rel) + rotational_motion_bound(pa, a, ma) + rotational_motion_bound(pb, b, mb) > threshold;
}

// Analytic round shapes (no reference counterpart).  The normal points
// from the collider towards the body; the point is the midpoint of the two
// surface points; penetration = rest offset - separation + 1e-5.
__device__ __forceinline__ V3 closest_on_segment(V3 p0, V3 p1, V3 q) {
    const V3 d = sub(p1, p0);
    const float dd = len2(d);
    const float t = dd > 0.0f ? clampf(dot(sub(q, p0), d) / dd, 0.0f, 1.0f) : 0.0f;
    return add(p0, mul(d, t));
}

__device__ bool point_shape_distance(V3 c, const PmhRigidBodyState &ss, const ShapeDev &sh, V3 &normal,
                                     V3 &surface, float &distance) {
    const Q o = sq(ss.orientation);
    const V3 sp = sv(ss.position);
    if (sh.kind == kShapeSphere) {
        const V3 d = sub(c, sp);
        normal = normalized_or(d, {0.0f, 1.0f, 0.0f});
        distance = len(d) - sh.sphere_radius;
        surface = add(sp, mul(normal, sh.sphere_radius));
        return true;
    }
    if (sh.kind == kShapeCapsule) {
        const V3 axis = rotate(o, {0.0f, sh.half_length, 0.0f});
        const V3 p = closest_on_segment(sub(sp, axis), add(sp, axis), c);
        const V3 d = sub(c, p);
        normal = normalized_or(d, rotate(o, {1.0f, 0.0f, 0.0f}));
        distance = len(d) - sh.sphere_radius;
        surface = add(p, mul(normal, sh.sphere_radius));
        return true;
    }
    if (!sh.is_box) return false;
    const V3 half = sarr(sh.box_half);
    const V3 l = inv_rotate(o, sub(c, sp));
    V3 q = {clampf(l.x, -half.x, half.x), clampf(l.y, -half.y, half.y), clampf(l.z, -half.z, half.z)};
    const bool inside = q.x == l.x && q.y == l.y && q.z == l.z;
    V3 ln;
    float dist;
    if (!inside) {
        const V3 d = sub(l, q);
        const float dl = len(d);
        ln = mul(d, 1.0f / dl);
        dist = dl;
    } else {
        const float dx = half.x - fabsf(l.x), dy = half.y - fabsf(l.y), dz = half.z - fabsf(l.z);
        if (dx <= dy && dx <= dz) { ln = {l.x < 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f}; q.x = l.x < 0.0f ? -half.x : half.x; dist = -dx; }
        else if (dy <= dz) { ln = {0.0f, l.y < 0.0f ? -1.0f : 1.0f, 0.0f}; q.y = l.y < 0.0f ? -half.y : half.y; dist = -dy; }
        else { ln = {0.0f, 0.0f, l.z < 0.0f ? -1.0f : 1.0f}; q.z = l.z < 0.0f ? -half.z : half.z; dist = -dz; }
    }
    normal = rotate(o, ln);
    surface = transform_point(ss, q);
    distance = dist;
    return true;
}

__device__ void push_round_contact(Manifold &m, V3 normal, V3 surface_collider, float gap, float margin,
                                   float reach) {
    if (gap > reach) return;
    const float rest = smin(margin, 1.0e-3f);
    Contact c;
    c.normal = normal;
    c.point = add(surface_collider, mul(normal, gap * 0.5f));
    c.pen = rest - gap + 1.0e-5f;
    c.impact_fraction = 1.0f;
    add_contact(m, c, smax(margin * 2.0f, 1.0e-4f));
}

__device__ bool analytic_manifold(const PmhRigidBodyState &as, const ShapeDev &am, const PmhRigidBodyState &bs,
                                  const ShapeDev &bm, float margin, float reach, Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const bool a_round = am.kind == kShapeSphere || am.kind == kShapeCapsule;
    const bool b_round = bm.kind == kShapeSphere || bm.kind == kShapeCapsule;
    if (!a_round && !b_round) return false;
    V3 n, s;
    float d;
    if (am.kind == kShapeSphere) {
        if (!point_shape_distance(sv(as.position), bs, bm, n, s, d)) return false;
        push_round_contact(out, n, s, d - am.sphere_radius, margin, reach);
        return true;
    }
    if (bm.kind == kShapeSphere) {
        if (!point_shape_distance(sv(bs.position), as, am, n, s, d)) return false;
        const V3 nb = mul(n, -1.0f);
        const V3 sb = sub(sv(bs.position), mul(n, bm.sphere_radius));
        push_round_contact(out, nb, sb, d - bm.sphere_radius, margin, reach);
        return true;
    }
    const bool a_cap = am.kind == kShapeCapsule;
    const ShapeDev &cap = a_cap ? am : bm;
    const PmhRigidBodyState &cs = a_cap ? as : bs;
    const ShapeDev &other = a_cap ? bm : am;
    const PmhRigidBodyState &os = a_cap ? bs : as;
    const V3 axis = rotate(sq(cs.orientation), {0.0f, cap.half_length, 0.0f});
    V3 ends[3] = {sub(sv(cs.position), axis), add(sv(cs.position), axis), sv(cs.position)};
    uint32_t samples = 2;
    if (other.kind == kShapeCapsule) {
        const V3 oaxis = rotate(sq(os.orientation), {0.0f, other.half_length, 0.0f});
        ends[2] = closest_on_segment(ends[0], ends[1],
                                     closest_on_segment(sub(sv(os.position), oaxis), add(sv(os.position), oaxis),
                                                        sv(cs.position)));
        samples = 3;
    }
    for (uint32_t k = 0; k < samples; ++k) {
        if (!point_shape_distance(ends[k], os, other, n, s, d)) return false;
        const float gap = d - cap.sphere_radius;
        if (a_cap) push_round_contact(out, n, s, gap, margin, reach);
        else push_round_contact(out, mul(n, -1.0f), sub(ends[k], mul(n, cap.sphere_radius)), gap, margin, reach);
    }
    return true;
}

__global__ void __launch_bounds__(128) k_narrowphase(const uint32_t *npairs, const PmhBodyParameters *P,
                                                     const ShapeDev *shapes, const V3 *verts, const uint32_t *idx,
                                                     const float4 *planes, const PmhRigidBodyState *prev,
                                                     const PmhRigidBodyState *pred, const uint2 *ab,
                                                     uint32_t *m_count, uint32_t *m_start, uint32_t *m_face,
                                                     float4 *m_init_rel, float4 *c_np, float4 *c_pt, float4 *c_aa,
                                                     float4 *c_ab, float4 *c_err, float4 *c_lam, float4 *c_pen,
                                                     float4 *c_acc, float4 *c_rep, uint32_t *c_flags,
                                                     uint32_t *c_pair, uint32_t point_cap, StepCounters *ctr) {
    const uint32_t M = *npairs;
    uint32_t unsupported = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        ShapeView sa{shapes[pa.mesh_index], verts, idx, planes};
        ShapeView sb{shapes[pb.mesh_index], verts, idx, planes};
        const PmhRigidBodyState A0 = prev[p.x], A = pred[p.x], B0 = prev[p.y], B = pred[p.y];
        const float margin = pa.collision_margin + pb.collision_margin;
        const float cr = correction_reach(pa, sa.s) + correction_reach(pb, sb.s);
        Manifold man;
        man.count = 0;
        man.face_patch = 0;
        if (sa.s.convex_face && sb.s.convex_face) {
            const bool swept = requires_swept(A0, A, sa.s, B0, B, sb.s, margin);
            bool face = false;
            if (swept) face = swept_convex_face_manifold(A0, A, sa, B0, B, sb, margin, man);
            if (!face) {
                face = convex_face_manifold(A, sa, B, sb, margin, cr, man);
                if (swept && man.count == 0u) face = false;
            }
            if (!face) { man.count = 0; man.face_patch = 0; ++unsupported; }
        } else if (!analytic_manifold(A, sa.s, B, sb.s, margin, margin + cr, man)) {
            man.count = 0;
            man.face_patch = 0;
            ++unsupported;
        }
        const auto sup = [](const ShapeDev &s) { return s.kind != kShapeMesh || (s.has_planes && s.idx_count <= 96u); };
        const bool persistent = sup(sa.s) && (pb.motion != kDynamic || sup(sb.s));
        // Each manifold owns a contiguous range of the point arrays; where
        // that range lies does not affect any result.
        uint32_t start = 0;
        if (man.count) {
            start = atomicAdd(&ctr->point_alloc, man.count);
            if (start + man.count > point_cap) {
                atomicOr(&ctr->status, kStatusPointOverflow);
                man.count = 0;
            }
        }
        m_count[m] = man.count;
        m_start[m] = start;
        m_face[m] = man.face_patch;
        m_init_rel[m] = f4(sub(sv(A.position), sv(B.position)), 0.0f);
        for (uint32_t k = 0; k < man.count; ++k) {
            const Contact &c = man.c[k];
            // pm_finalize_rigid_contact_manifolds (geometry.slang:2805-2873)
            const V3 av = add(sv(A.linear_velocity), cross(sv(A.angular_velocity), sub(c.point, sv(A.position))));
            const V3 bv = add(sv(B.linear_velocity), cross(sv(B.angular_velocity), sub(c.point, sv(B.position))));
            const float speed = dot(sub(av, bv), c.normal);
            const uint32_t i = start + k;
            const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            c_np[i] = f4(c.normal, c.pen);
            c_pt[i] = f4(c.point, speed);
            c_aa[i] = make_float4(0.0f, 0.0f, 0.0f, c.impact_fraction);   // w: impact fraction
            c_ab[i] = z;
            c_err[i] = z;
            c_lam[i] = z;
            c_pen[i] = z;
            c_acc[i] = z;
            c_rep[i] = z;
            c_flags[i] = persistent ? 2u : 0u;
            c_pair[i] = m;
        }
    }
    if (unsupported) atomicAdd(&ctr->unsupported, unsupported);
}

__global__ void k_nonempty(const uint32_t *npairs, const uint32_t *m_count, uint32_t *flag) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x)
        flag[m] = m_count[m] != 0u ? 1u : 0u;
}

// An overflowing step solves nothing (its manifolds are incomplete) and is
// reported as PM_ERROR_CAPACITY_EXCEEDED.
__global__ void k_guard_points(StepCounters *ctr, uint32_t point_cap) {
    if (threadIdx.x != 0) return;
    const uint32_t P = ctr->point_alloc;
    if (P > point_cap) {
        atomicOr(&ctr->status, kStatusPointOverflow);
        ctr->points = 0;
        ctr->pairs = 0;
    } else {
        ctr->points = P;
    }
}

// pm_load_rigid_contact_cache (geometry.slang:2941-3009) against the
// previous step's saved pairs, found by binary search on their sorted codes.
__global__ void k_cache_load(const uint32_t *npairs, const uint32_t *code, const uint2 *ab, const uint32_t *m_count,
                             const uint32_t *m_start, const PmhRigidBodyState *pred, const uint32_t *saved_code,
                             const uint32_t *saved_count, const float4 *saved, const uint32_t *n_saved_ptr,
                             float4 *c_np, float4 *c_pt, float4 *c_aa, float4 *c_ab, float4 *c_lam,
                             float4 *c_pen, float4 *c_acc, uint32_t *c_flags, float dt, StepCounters *ctr) {
    const uint32_t M = *npairs;
    const uint32_t ns = *n_saved_ptr;
    uint32_t hits = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint32_t key = code[m];
        uint32_t lo = 0, hi = ns;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (saved_code[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (lo >= ns || saved_code[lo] != key) continue;
        const uint32_t sc = saved_count[lo];
        const float4 *rec = saved + (size_t)lo * 8 * kCacheFloat4PerPoint;
        const PmhRigidBodyState body = pred[ab[m].x];
        uint32_t used = 0;
        for (uint32_t k = 0; k < m_count[m]; ++k) {
            const uint32_t c = m_start[m] + k;
            const V3 n = v3(c_np[c]), point = v3(c_pt[c]);
            const V3 local = inv_rotate(sq(body.orientation), sub(point, sv(body.position)));
            float nearest = 0.02f * 0.02f;
            uint32_t match = 8;
            for (uint32_t q = 0; q < sc; ++q) {
                const float4 *r = rec + q * kCacheFloat4PerPoint;
                if ((used & (1u << q)) != 0u || dot(n, v3(r[1])) < 0.99f) continue;
                const float d = len2(sub(local, v3(r[0])));
                if (d < nearest) { nearest = d; match = q; }
            }
            if (match == 8u) continue;
            used |= 1u << match;
            const float4 *r = rec + match * kCacheFloat4PerPoint;
            // r0: local point, normal impulse; r1: normal; r2: friction
            // impulse; r3: lambdas; r4: penalties; r5: anchor a; anchor b in
            // r1.w/r2.w/r4.w
            V3 t0, t1;
            contact_basis(n, t0, t1);
            const V3 fi = v3(r[2]);
            const float4 lam = r[3], pen = r[4];
            c_lam[c] = make_float4(lam.x, dot(fi, t0) / dt, dot(fi, t1) / dt, 0.0f);
            c_pen[c] = make_float4(pen.x, pen.y, pen.z, 0.0f);
            c_aa[c] = f4(v3(r[5]), c_aa[c].w);
            c_ab[c] = make_float4(r[1].w, r[2].w, r[4].w, 0.0f);
            c_flags[c] |= 1u;
            const V3 af = sub(fi, mul(n, dot(fi, n)));
            c_acc[c] = make_float4(r[0].w, af.x, af.y, af.z);
            ++hits;
        }
    }
    if (hits) atomicAdd(&ctr->cache_hits, hits);
}

// pm_save_rigid_contact_cache (geometry.slang:3011-3070): the pair of
// response slot s (its rank among non-empty manifolds) is saved at s.
__global__ void k_cache_save(const uint32_t *npairs, const uint32_t *code, const uint2 *ab, const uint32_t *m_count,
                             const uint32_t *m_start, const uint32_t *m_slot, const PmhRigidBodyState *states,
                             const float4 *c_np, const float4 *c_pt, const float4 *c_aa, const float4 *c_ab,
                             const float4 *c_lam, const float4 *c_pen, const float4 *c_acc, uint32_t *saved_code,
                             uint32_t *saved_count, float4 *saved, uint32_t *n_saved, uint32_t cap,
                             const StepCounters *ctr) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint32_t s = m_slot[m];
        if (s >= cap) continue;
        saved_code[s] = code[m];
        saved_count[s] = m_count[m];
        const PmhRigidBodyState body = states[ab[m].x];
        float4 *rec = saved + (size_t)s * 8 * kCacheFloat4PerPoint;
        for (uint32_t k = 0; k < m_count[m]; ++k) {
            const uint32_t c = m_start[m] + k;
            const V3 local = inv_rotate(sq(body.orientation), sub(v3(c_pt[c]), sv(body.position)));
            const float4 acc = c_acc[c], bb = c_ab[c];
            float4 *r = rec + k * kCacheFloat4PerPoint;
            r[0] = f4(local, acc.x);
            r[1] = f4(v3(c_np[c]), bb.x);
            r[2] = make_float4(acc.y, acc.z, acc.w, bb.y);
            r[3] = c_lam[c];
            r[4] = make_float4(c_pen[c].x, c_pen[c].y, c_pen[c].z, bb.z);
            r[5] = f4(v3(c_aa[c]), 0.0f);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_saved = std::min(ctr->manifolds, cap);
}

__global__ void k_clamp_speeds(uint32_t n, const PmhBodyParameters *P, PmhRigidBodyState *s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || P[i].motion != kDynamic) return;
    s[i].linear_velocity = pv(clamp_length(sv(s[i].linear_velocity), P[i].maximum_linear_speed));
    s[i].angular_velocity = pv(clamp_length(sv(s[i].angular_velocity), P[i].maximum_angular_speed));
}

__global__ void k_step_begin(StepCounters *ctr) {
    if (threadIdx.x != 0) return;
    ctr->pairs = 0;
    ctr->points = 0;
    ctr->manifolds = 0;
    ctr->colors = 0;
    ctr->islands = 0;
    ctr->impact = 0;
    ctr->unsupported = 0;
    ctr->cache_hits = 0;
    ctr->large_count = 0;
    ctr->color_ticket = 0;
    ctr->point_alloc = 0;
    ctr->cell_size_bits = 0.0f;
}

static uint32_t blocks_for(uint32_t items, uint32_t block = 256, uint32_t cap = 4096) {
    uint32_t g = (items + block - 1) / block;
    return std::max<uint32_t>(1, std::min(g, cap));
}

PmhResult World::step_once(uint32_t k_step) {
    hipStream_t s = stream;
    const uint32_t n = body_count;
    Solver &v = solver;
    StepCounters *ctr = v.counters.p;
    const uint32_t bg = blocks_for(n);
    const uint32_t pg = blocks_for(v.pair_cap, 256, 2048);
    auto mark = [&](int k) { if (timing_enabled) (void)hipEventRecord(events[k], s); };
    mark(0);
    k_step_begin<<<1, 64, 0, s>>>(ctr);
    PmhRigidBodyState *cur = state_cur.p, *prev = state_prev.p, *pred = state_next.p;
    k_integrate_proxy<<<bg, 256, 0, s>>>(n, params.p, shapes_dev.p, cur, prev, pred, proxy_mn.p, proxy_mx.p,
                                         gravity, dt, ctr);
    mark(1);
    // Broadphase.
    PMH_CHECK(hipMemsetAsync(bucket_count.p, 0, sizeof(uint32_t) * (table_size + 1), s));
    k_bp_cells<<<bg, 256, 0, s>>>(n, proxy_mn.p, proxy_mx.p, cell_of.p, bucket_count.p, table_size - 1,
                                  large.p, ctr, grid_cell);
    PmhResult r = exclusive_scan(s, bucket_count.p, bucket_start.p, table_size, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    PMH_CHECK(hipMemsetAsync(bucket_count.p, 0, sizeof(uint32_t) * table_size, s));
    k_bp_scatter<<<bg, 256, 0, s>>>(n, cell_of.p, bucket_start.p, bucket_count.p, bucket_entries.p);
    k_bp_query<0><<<blocks_for(n, 64), 64, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, cell_of.p,
                                                  bucket_start.p, bucket_entries.p, table_size - 1, large.p,
                                                  cand_count.p, nullptr, nullptr, nullptr, ctr);
    k_bp_large<0><<<64, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, large.p, cand_count.p, nullptr, nullptr,
                                     nullptr, ctr);
    r = exclusive_scan(s, cand_count.p, cand_offset.p, n, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    k_set_pairs<<<1, 64, 0, s>>>(ctr, cand_offset.p, n, v.pair_cap);
    k_bp_large<1><<<64, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, large.p, cand_count.p, cand_offset.p,
                                     v.code.p, v.ab.p, ctr);
    k_bp_query<1><<<blocks_for(n, 64), 64, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, cell_of.p,
                                                  bucket_start.p, bucket_entries.p, table_size - 1, large.p,
                                                  cand_count.p, cand_offset.p, v.code.p, v.ab.p, ctr);
    mark(2);
    // Narrowphase + point compaction.
    k_narrowphase<<<blocks_for(v.pair_cap, 128, 4096), 128, 0, s>>>(
        &ctr->pairs, params.p, shapes_dev.p, verts_dev.p, idx_dev.p, planes_dev.p, prev, pred, v.ab.p,
        v.m_count.p, v.m_start.p, v.m_face.p, v.m_init_rel.p, v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p,
        v.c_err.p, v.c_lam.p, v.c_pen.p, v.c_acc.p, v.c_rep.p, v.c_flags.p, v.c_pair.p, v.point_cap, ctr);
    k_guard_points<<<1, 64, 0, s>>>(ctr, v.point_cap);
    mark(3);
    // Response slots and the contact cache.
    k_nonempty<<<pg, 256, 0, s>>>(&ctr->pairs, v.m_count.p, v.m_impact.p);
    r = exclusive_scan(s, v.m_impact.p, v.m_slot.p, v.pair_cap, &ctr->pairs, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    const int rd = cache_parity, wr = cache_parity ^ 1;
    k_cache_load<<<pg, 256, 0, s>>>(&ctr->pairs, v.code.p, v.ab.p, v.m_count.p, v.m_start.p, pred,
                                    cache_code[rd].p, cache_count[rd].p, cache_data[rd].p, &cache_n.p[rd],
                                    v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_lam.p, v.c_pen.p, v.c_acc.p,
                                    v.c_flags.p, dt, ctr);
    mark(4);
    SolveParams sp{dt, gravity, base_iterations, requested_iterations, profile, cooperative ? 1u : 0u, solve_lanes};
    SolveTiming st;
    if (timing_enabled) {
        st.start = solve_events[2 * k_step];
        st.stop = solve_events[2 * k_step + 1];
    }
    r = avbd_prepare_solve(s, v, params.p, prev, pred, sp, timing_enabled ? &st : nullptr);
    solve_timing.launches += st.launches;
    if (r != PMH_SUCCESS) return r;
    mark(5);
    k_clamp_speeds<<<bg, 256, 0, s>>>(n, params.p, pred);
    k_cache_save<<<pg, 256, 0, s>>>(&ctr->pairs, v.code.p, v.ab.p, v.m_count.p, v.m_start.p, v.m_slot.p, pred,
                                    v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_lam.p, v.c_pen.p, v.c_acc.p,
                                    cache_code[wr].p, cache_count[wr].p, cache_data[wr].p, &cache_n.p[wr],
                                    cache_capacity, ctr);
    mark(6);
    cache_parity = wr;
    std::swap(state_cur.p, state_next.p);
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

}  // namespace pmh
