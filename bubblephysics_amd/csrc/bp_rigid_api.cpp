// bp_rigid_api.cpp -- host side of the rigid solver for worlds without bubbles
// (bp_rigid_* in include/bubblephysics.h).
//
// A bp_rigid_batch packs every world into one device arena (bodies on two sides, geometry,
// segments, joint tables, rows and corrections at a per-world stride) and advances all worlds
// with one launch of rigid_step_kernel (bp_rigid.hip), one workgroup per world, every step of the
// call inside the launch. The host keeps each world's input; the device keeps the published
// bodies between calls. Joint tables are built on the host as NativeJointState.prepare builds them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "bp_batch.h"
#include "bp_rigid.h"

using bph::set_error;

namespace {

struct HostRigidWorld {
    bp_rigid_config config{};
    std::vector<bp_rigid_body> bodies;
    std::vector<float> vertices;       // float2 pairs
    std::vector<uint32_t> triangles;   // uint4 records
    std::vector<bp_segment> segments;
    std::vector<bp_joint> joints;
    std::vector<uint32_t> references;  // (joint, side) pairs
    std::vector<uint32_t> ranges;      // uint4 per body (empty: no joint tables)
    std::vector<uint64_t> partners;
    bool failed = false;
    bool lastOk = false;
    bool stepped = false;
};

uint32_t row_stride(const HostRigidWorld &w) {
    // A row holds at most one entry per other body, so a stride of min(k, n - 1) keeps every entry
    // the reference's k-entry row would keep.
    const uint32_t n = (uint32_t)w.bodies.size();
    return std::max(1u, std::min(w.config.candidate_limit, n > 1 ? n - 1 : 1u));
}

bool finite_n(const float *v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

bp_status validate(HostRigidWorld &w) {
    bp_rigid_config &c = w.config;
    if (!(c.time_step > 0.0f) || !std::isfinite(c.time_step)) { set_error("time_step must be finite and positive"); return BP_ERR_INVALID_SNAPSHOT; }
    if (!finite_n(c.region, 4) || !(c.region[2] > c.region[0]) || !(c.region[3] > c.region[1])) {
        set_error("region must be finite with max > min");
        return BP_ERR_INVALID_SNAPSHOT;
    }
    if (c.candidate_limit == 0) c.candidate_limit = 128;
    if (c.candidate_limit > 128) { set_error("candidate_limit must be 1..128"); return BP_ERR_INVALID_ARGUMENT; }
    if (c.iterations == 0) c.iterations = 8;
    if (c.iterations > 1024) { set_error("iterations must be at most 1024"); return BP_ERR_INVALID_ARGUMENT; }
    if (c.motion_prelude && !finite_n(c.global_force, 2)) { set_error("global_force must be finite"); return BP_ERR_INVALID_SNAPSHOT; }
    const size_t nv = w.vertices.size() / 2, nt = w.triangles.size() / 4, n = w.bodies.size();
    if (n > 65536) { set_error("at most 65,536 bodies per rigid world"); return BP_ERR_CAPACITY_OVERFLOW; }
    for (size_t i = 0; i < n; ++i) {
        const bp_rigid_body &b = w.bodies[i];
        if (i > 0 && !(w.bodies[i - 1].id < b.id)) { set_error("bodies must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
        if ((uint64_t)b.ranges[0] + b.ranges[1] > nv || (uint64_t)b.ranges[2] + b.ranges[3] > nt) {
            set_error("body geometry range outside the vertex or triangle arrays");
            return BP_ERR_INVALID_SNAPSHOT;
        }
        if (b.ranges[1] == 0) { set_error("body without vertices"); return BP_ERR_INVALID_SNAPSHOT; }
        const float kind = b.shape[0];
        if (!(kind == 0.0f || kind == 1.0f || kind == 2.0f)) { set_error("unknown body shape"); return BP_ERR_INVALID_SNAPSHOT; }
        if (kind == 1.0f && b.ranges[1] != 1) { set_error("a circle has one core vertex"); return BP_ERR_INVALID_SNAPSHOT; }
        if (kind == 2.0f && b.ranges[1] != 2) { set_error("a capsule has two core vertices"); return BP_ERR_INVALID_SNAPSHOT; }
        for (uint32_t t = 0; t < b.ranges[3]; ++t)
            for (int k = 0; k < 3; ++k)
                if (w.triangles[4 * ((size_t)b.ranges[2] + t) + k] >= b.ranges[1]) {
                    set_error("triangle index outside its body's vertices");
                    return BP_ERR_INVALID_SNAPSHOT;
                }
        if (!finite_n(b.pose, 4) || !finite_n(b.previous, 4) || !finite_n(b.motion, 4) || !finite_n(b.physical, 4) ||
            !finite_n(b.shape, 4) || !finite_n(b.limits, 4)) {
            set_error("body " + std::to_string(b.id) + " has a non-finite field");
            return BP_ERR_INVALID_SNAPSHOT;
        }
        const bool dynamic = !b.flags[0] && !b.filter[2];
        if (c.motion_prelude && dynamic && !(b.mass > 0.0f && std::isfinite(b.mass))) {
            set_error("the motion prelude needs a positive mass on every dynamic body");
            return BP_ERR_INVALID_SNAPSHOT;
        }
    }
    for (size_t v = 0; v < w.vertices.size(); ++v)
        if (!std::isfinite(w.vertices[v])) { set_error("non-finite vertex"); return BP_ERR_INVALID_SNAPSHOT; }
    for (size_t i = 1; i < w.segments.size(); ++i)
        if (!(w.segments[i - 1].id < w.segments[i].id)) { set_error("segments must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
    if (!w.joints.empty() && w.ranges.empty()) { set_error("joints need their tables (bp_rigid_joint_tables)"); return BP_ERR_INVALID_ARGUMENT; }
    if (!w.ranges.empty() && w.ranges.size() != 4 * n) { set_error("joint ranges must hold one uint4 per body"); return BP_ERR_INVALID_ARGUMENT; }
    for (const bp_joint &j : w.joints) {
        if ((j.ends[0] != BP_JOINT_WORLD_END && j.ends[0] >= n) || j.ends[1] >= n || j.ends[2] > BP_JOINT_WELD) {
            set_error("joint end outside the world's bodies, or unknown joint kind");
            return BP_ERR_INVALID_SNAPSHOT;
        }
    }
    const size_t nrefs = w.references.size() / 2;
    for (size_t r = 0; r < nrefs; ++r)
        if (w.references[2 * r] >= w.joints.size() || w.references[2 * r + 1] > 1) {
            set_error("joint reference outside the joints");
            return BP_ERR_INVALID_SNAPSHOT;
        }
    for (size_t i = 0; i < (w.ranges.empty() ? 0 : n); ++i) {
        const uint32_t *r = &w.ranges[4 * i];
        if ((uint64_t)r[0] + r[1] > nrefs || (uint64_t)r[2] + r[3] > w.partners.size()) {
            set_error("joint range outside the reference or partner arrays");
            return BP_ERR_INVALID_SNAPSHOT;
        }
    }
    return BP_OK;
}

template <class T>
bp_status grow(T *&p, size_t &cap, size_t need) {
    need = std::max<size_t>(need, 1);
    if (cap >= need) return BP_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = need + need / 4;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), want * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return BP_ERR_ALLOCATION_FAILED;
    }
    cap = want;
    return BP_OK;
}

}  // namespace

struct bp_rigid_batch {
    std::vector<HostRigidWorld> worlds;
    bool packed = false;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    float lastMs = 0.0f;

    bpr::RigidArena arena{};
    bpr::RigidDesc *descs = nullptr;
    std::vector<bpr::RigidDesc> hostDescs;
    bp_rigid_frame_info *infos = nullptr;
    size_t capBodies[2] = {0, 0}, capVerts = 0, capTris = 0, capSegs = 0, capJoints = 0, capJointWork = 0, capJcorr = 0,
           capRefs = 0, capRanges = 0, capPartners = 0, capRows = 0, capCorr = 0, capSlots = 0, capRowCount = 0,
           capRowStart = 0, capWalls = 0, capDescs = 0, capInfos = 0, capTwin = 0, capToi = 0, capToiWork = 0;
    bp_rigid_body *bodies[2] = {nullptr, nullptr};
    float2 *verts = nullptr;
    uint4 *tris = nullptr;
    bp_segment *segs = nullptr;
    bp_joint *jointInput = nullptr, *joints = nullptr;
    bpr::RigidCorr *jcorr = nullptr, *corr = nullptr;
    uint2 *refs = nullptr;
    uint4 *jranges = nullptr;
    uint64_t *partners = nullptr;
    bpr::RigidCand *rows = nullptr;
    uint32_t *slots = nullptr, *rowCount = nullptr, *rowStart = nullptr, *twin = nullptr;
    float *toi = nullptr;
    uint32_t *toiWork = nullptr;   // per used slot: impact-time slices, their prefix, the first hit
    uint32_t *orders = nullptr;
    size_t capOrders = 0;
    uint32_t smallCount = 0, largeCount = 0;
    bpr::RigidWall *walls = nullptr;

    ~bp_rigid_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        void *ptrs[] = {bodies[0], bodies[1], verts, tris, segs, jointInput, joints, jcorr, corr, refs, jranges,
                        partners, rows, slots, rowCount, rowStart, walls, descs, infos, twin, toi, toiWork, orders};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
        if (stream) (void)hipStreamDestroy(stream);
    }

    bp_status ensure_stream() {
        if (stream) return BP_OK;
        BP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BP_HIP(hipEventCreate(&start));
        BP_HIP(hipEventCreate(&stop));
        return BP_OK;
    }

    // Reads every world's published bodies and flags back into the host copies (both body sides
    // in two copies; each world takes its published side).
    bp_status sync_from_device() {
        if (!packed) return BP_OK;
        BP_HIP(hipMemcpyAsync(hostDescs.data(), descs, hostDescs.size() * sizeof(bpr::RigidDesc), hipMemcpyDeviceToHost, stream));
        size_t total = 0;
        for (const HostRigidWorld &w : worlds) total += w.bodies.size();
        std::vector<bp_rigid_body> side[2];
        for (int k = 0; k < 2; ++k) {
            side[k].resize(total);
            if (total) BP_HIP(hipMemcpyAsync(side[k].data(), bodies[k], total * sizeof(bp_rigid_body), hipMemcpyDeviceToHost, stream));
        }
        BP_HIP(hipStreamSynchronize(stream));
        for (size_t wi = 0; wi < worlds.size(); ++wi) {
            HostRigidWorld &w = worlds[wi];
            const bpr::RigidDesc &d = hostDescs[wi];
            std::copy(side[d.side].begin() + d.bodyOffset, side[d.side].begin() + d.bodyOffset + w.bodies.size(), w.bodies.begin());
            w.failed = d.failed != 0;
            w.lastOk = d.lastOk != 0;
        }
        return BP_OK;
    }

    bp_status pack() {
        bp_status s = ensure_stream();
        if (s != BP_OK) return s;
        size_t nb = 0, nv = 0, nt = 0, ns = 0, nj = 0, nr = 0, np = 0, nrows = 0;
        hostDescs.assign(worlds.size(), bpr::RigidDesc{});
        for (size_t wi = 0; wi < worlds.size(); ++wi) {
            const HostRigidWorld &w = worlds[wi];
            bpr::RigidDesc &d = hostDescs[wi];
            d.bodyOffset = (uint32_t)nb;
            d.bodyCount = (uint32_t)w.bodies.size();
            d.vertexOffset = (uint32_t)nv;
            d.triangleOffset = (uint32_t)nt;
            d.segmentOffset = (uint32_t)ns;
            d.segmentCount = (uint32_t)w.segments.size();
            d.jointOffset = (uint32_t)nj;
            d.jointCount = (uint32_t)w.joints.size();
            d.refOffset = (uint32_t)nr;
            d.partnerOffset = (uint32_t)np;
            d.hasJointTables = w.ranges.empty() ? 0u : 1u;
            d.rowOffset = (uint32_t)nrows;
            d.stride = row_stride(w);
            d.k = w.config.candidate_limit;
            d.iterations = w.config.iterations;
            d.prelude = w.config.motion_prelude ? 1u : 0u;
            d.dt = w.config.time_step;
            d.force[0] = w.config.global_force[0];
            d.force[1] = w.config.global_force[1];
            std::memcpy(d.region, w.config.region, sizeof(d.region));
            d.side = 0;
            d.failed = w.failed ? 1u : 0u;
            d.lastOk = w.lastOk ? 1u : 0u;
            nb += w.bodies.size();
            nv += w.vertices.size() / 2;
            nt += w.triangles.size() / 4;
            ns += w.segments.size();
            nj += w.joints.size();
            nr += w.references.size() / 2;
            np += w.partners.size();
            nrows += (size_t)d.stride * w.bodies.size();
        }
        if (nb > 0xffffffffull || nrows > 0xffffffffull) { set_error("rigid batch too large"); return BP_ERR_CAPACITY_OVERFLOW; }
        bp_status g = BP_OK;
        auto G = [&](bp_status x) { if (g == BP_OK) g = x; };
        G(grow(bodies[0], capBodies[0], nb));
        G(grow(bodies[1], capBodies[1], nb));
        G(grow(verts, capVerts, nv));
        G(grow(tris, capTris, nt));
        G(grow(segs, capSegs, ns));
        G(grow(jointInput, capJoints, nj));
        G(grow(joints, capJointWork, nj));
        G(grow(jcorr, capJcorr, nj));
        G(grow(refs, capRefs, nr));
        G(grow(jranges, capRanges, nb));
        G(grow(partners, capPartners, np));
        G(grow(rows, capRows, nrows));
        G(grow(corr, capCorr, nrows));
        G(grow(slots, capSlots, nrows));
        G(grow(twin, capTwin, nrows));
        G(grow(toi, capToi, nrows));
        G(grow(toiWork, capToiWork, 3 * nrows));
        G(grow(rowCount, capRowCount, nb));
        G(grow(rowStart, capRowStart, nb));
        G(grow(walls, capWalls, nb));
        G(grow(descs, capDescs, worlds.size()));
        G(grow(orders, capOrders, worlds.size()));
        if (g != BP_OK) return g;
        // Every array staged contiguously on the host and uploaded with one copy (a copy per world and
        // array would be tens of thousands of small transfers for a large batch).
        std::vector<bp_rigid_body> hb;
        std::vector<float> hv;
        std::vector<uint32_t> ht, hr, hj;
        std::vector<bp_segment> hs;
        std::vector<bp_joint> hjo;
        std::vector<uint64_t> hp;
        hb.reserve(nb);
        hv.reserve(2 * nv);
        ht.reserve(4 * nt);
        hs.reserve(ns);
        hjo.reserve(nj);
        hr.reserve(2 * nr);
        hj.assign(4 * nb, 0u);
        hp.reserve(np);
        for (size_t wi = 0; wi < worlds.size(); ++wi) {
            const HostRigidWorld &w = worlds[wi];
            hb.insert(hb.end(), w.bodies.begin(), w.bodies.end());
            hv.insert(hv.end(), w.vertices.begin(), w.vertices.end());
            ht.insert(ht.end(), w.triangles.begin(), w.triangles.end());
            hs.insert(hs.end(), w.segments.begin(), w.segments.end());
            hjo.insert(hjo.end(), w.joints.begin(), w.joints.end());
            hr.insert(hr.end(), w.references.begin(), w.references.end());
            if (!w.ranges.empty()) std::copy(w.ranges.begin(), w.ranges.end(), hj.begin() + 4 * (size_t)hostDescs[wi].bodyOffset);
            hp.insert(hp.end(), w.partners.begin(), w.partners.end());
        }
        auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
            return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
        };
        BP_HIP(up(bodies[0], hb.data(), hb.size() * sizeof(bp_rigid_body)));
        BP_HIP(up(verts, hv.data(), hv.size() * sizeof(float)));
        BP_HIP(up(tris, ht.data(), ht.size() * sizeof(uint32_t)));
        BP_HIP(up(segs, hs.data(), hs.size() * sizeof(bp_segment)));
        BP_HIP(up(jointInput, hjo.data(), hjo.size() * sizeof(bp_joint)));
        BP_HIP(up(joints, hjo.data(), hjo.size() * sizeof(bp_joint)));
        BP_HIP(up(refs, hr.data(), hr.size() * sizeof(uint32_t)));
        BP_HIP(up(jranges, hj.data(), hj.size() * sizeof(uint32_t)));
        BP_HIP(up(partners, hp.data(), hp.size() * sizeof(uint64_t)));
        if (nb) BP_HIP(hipMemsetAsync(rowCount, 0, nb * sizeof(uint32_t), stream));
        if (nb) BP_HIP(hipMemsetAsync(walls, 0, nb * sizeof(bpr::RigidWall), stream));
        if (nrows) BP_HIP(hipMemsetAsync(corr, 0, nrows * sizeof(bpr::RigidCorr), stream));
        if (!hostDescs.empty())
            BP_HIP(hipMemcpyAsync(descs, hostDescs.data(), hostDescs.size() * sizeof(bpr::RigidDesc), hipMemcpyHostToDevice, stream));
        // Launch lists: worlds that fit the LDS kernel first, then the others.
        std::vector<uint32_t> order;
        order.reserve(worlds.size());
        for (size_t wi = 0; wi < worlds.size(); ++wi)
            if (worlds[wi].bodies.size() <= bpr::kRigidLdsBodies) order.push_back((uint32_t)wi);
        smallCount = (uint32_t)order.size();
        for (size_t wi = 0; wi < worlds.size(); ++wi)
            if (worlds[wi].bodies.size() > bpr::kRigidLdsBodies) order.push_back((uint32_t)wi);
        largeCount = (uint32_t)(order.size() - smallCount);
        if (!order.empty())
            BP_HIP(hipMemcpyAsync(orders, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        BP_HIP(hipStreamSynchronize(stream));
        arena.bodies[0] = bodies[0];
        arena.bodies[1] = bodies[1];
        arena.verts = verts;
        arena.tris = tris;
        arena.segs = segs;
        arena.jointInput = jointInput;
        arena.joints = joints;
        arena.jcorr = jcorr;
        arena.refs = refs;
        arena.jranges = jranges;
        arena.partners = partners;
        arena.rows = rows;
        arena.corr = corr;
        arena.slots = slots;
        arena.twin = twin;
        arena.toi = toi;
        arena.toiSteps = toiWork;
        arena.toiStart = toiWork + nrows;
        arena.toiFirst = toiWork + 2 * nrows;
        arena.rowCount = rowCount;
        arena.rowStart = rowStart;
        arena.walls = walls;
        packed = true;
        return BP_OK;
    }

    bp_status step(uint32_t steps, bp_rigid_frame_info *out) {
        if (steps == 0) return BP_OK;
        if (!packed) {
            bp_status s = pack();
            if (s != BP_OK) return s;
        }
        const size_t W = worlds.size();
        if (W == 0) return BP_OK;
        bp_status s = grow(infos, capInfos, W * steps);
        if (s != BP_OK) return s;
        BP_HIP(hipEventRecord(start, stream));
        hipError_t e = bp_launch_rigid_step(&arena, descs, orders, smallCount, 1, infos, (uint32_t)W, steps, stream);
        if (e == hipSuccess)
            e = bp_launch_rigid_step(&arena, descs, orders + smallCount, largeCount, 0, infos, (uint32_t)W, steps, stream);
        if (e != hipSuccess) { set_error(std::string("rigid kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
        BP_HIP(hipEventRecord(stop, stream));
        BP_HIP(hipMemcpyAsync(hostDescs.data(), descs, W * sizeof(bpr::RigidDesc), hipMemcpyDeviceToHost, stream));
        if (out) BP_HIP(hipMemcpyAsync(out, infos, W * steps * sizeof(bp_rigid_frame_info), hipMemcpyDeviceToHost, stream));
        BP_HIP(hipStreamSynchronize(stream));
        BP_HIP(hipEventElapsedTime(&lastMs, start, stop));
        for (size_t wi = 0; wi < W; ++wi) {
            worlds[wi].failed = hostDescs[wi].failed != 0;
            worlds[wi].lastOk = hostDescs[wi].lastOk != 0;
            worlds[wi].stepped = true;
        }
        return BP_OK;
    }

    bp_status read_bodies(uint32_t wi, bp_rigid_body *out, uint32_t count) {
        if (wi >= worlds.size() || !out) return BP_ERR_INVALID_ARGUMENT;
        HostRigidWorld &w = worlds[wi];
        if (count != w.bodies.size()) { set_error("count must equal the world's body count"); return BP_ERR_INVALID_ARGUMENT; }
        if (!packed) {
            std::copy(w.bodies.begin(), w.bodies.end(), out);
            return BP_OK;
        }
        const bpr::RigidDesc &d = hostDescs[wi];
        if (count) BP_HIP(hipMemcpyAsync(out, bodies[d.side] + d.bodyOffset, count * sizeof(bp_rigid_body), hipMemcpyDeviceToHost, stream));
        BP_HIP(hipStreamSynchronize(stream));
        return BP_OK;
    }

    bp_status read_joints(uint32_t wi, bp_joint *out, uint32_t count) {
        if (wi >= worlds.size() || (!out && count)) return BP_ERR_INVALID_ARGUMENT;
        HostRigidWorld &w = worlds[wi];
        if (count != w.joints.size()) { set_error("count must equal the world's joint count"); return BP_ERR_INVALID_ARGUMENT; }
        if (!packed || !w.stepped) {
            std::copy(w.joints.begin(), w.joints.end(), out);
            return BP_OK;
        }
        const bpr::RigidDesc &d = hostDescs[wi];
        if (count) BP_HIP(hipMemcpyAsync(out, joints + d.jointOffset, count * sizeof(bp_joint), hipMemcpyDeviceToHost, stream));
        BP_HIP(hipStreamSynchronize(stream));
        return BP_OK;
    }

    // NativeRigidPolygonState.contactRecords (NativeRigidPolygonState.swift:367-428), body part:
    // the body-pair entries of each body's row that it owns (lower index), then the wall records.
    bp_status contacts(uint32_t wi, bp_contact *out, uint32_t capacity, uint32_t *outCount) {
        if (wi >= worlds.size() || !outCount || (capacity && !out)) return BP_ERR_INVALID_ARGUMENT;
        HostRigidWorld &w = worlds[wi];
        *outCount = 0;
        if (!packed || !w.stepped || !w.lastOk) {
            set_error("the world's last step did not publish: no contact records");
            return BP_ERR_UNAVAILABLE;
        }
        const bpr::RigidDesc &d = hostDescs[wi];
        const uint32_t n = d.bodyCount, stride = d.stride;
        std::vector<uint32_t> counts(n);
        std::vector<bpr::RigidCand> keys((size_t)n * stride);
        std::vector<bpr::RigidCorr> cs((size_t)n * stride);
        std::vector<bpr::RigidWall> ws(n);
        std::vector<bp_rigid_body> bs(n);
        if (n) {
            BP_HIP(hipMemcpyAsync(counts.data(), rowCount + d.bodyOffset, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
            BP_HIP(hipMemcpyAsync(keys.data(), rows + d.rowOffset, keys.size() * sizeof(bpr::RigidCand), hipMemcpyDeviceToHost, stream));
            BP_HIP(hipMemcpyAsync(cs.data(), corr + d.rowOffset, cs.size() * sizeof(bpr::RigidCorr), hipMemcpyDeviceToHost, stream));
            BP_HIP(hipMemcpyAsync(ws.data(), walls + d.bodyOffset, n * sizeof(bpr::RigidWall), hipMemcpyDeviceToHost, stream));
            BP_HIP(hipMemcpyAsync(bs.data(), bodies[d.side] + d.bodyOffset, n * sizeof(bp_rigid_body), hipMemcpyDeviceToHost, stream));
        }
        BP_HIP(hipStreamSynchronize(stream));
        uint32_t count = 0;
        auto push = [&](const bp_contact &c) {
            if (count < capacity) out[count] = c;
            ++count;
        };
        std::set<std::pair<uint64_t, uint64_t>> emitted;
        for (uint32_t body = 0; body < n; ++body)
            for (uint32_t slot = 0; slot < counts[body]; ++slot) {
                const size_t at = (size_t)body * stride + slot;
                const bpr::RigidCand &k = keys[at];
                if (k.kind == 3u && k.low != body) continue;
                const float nx = cs[at].first[3], ny = cs[at].second[3];
                if (!(nx != 0.0f || ny != 0.0f)) continue;
                bp_contact r{};
                r.first = k.first;
                r.second = k.second;
                r.kind = BP_CONTACT_RIGID;
                r.normal[0] = nx;
                r.normal[1] = ny;
                r.point[0] = cs[at].vfirst[3];
                r.point[1] = cs[at].vsecond[3];
                r.impulse = k.impulse;
                r.index_b = k.high;
                emitted.insert({(uint64_t)k.first, (uint64_t)k.second});
                push(r);
            }
        std::map<uint64_t, uint32_t> position;
        for (uint32_t i = 0; i < n; ++i) position[bs[i].id] = i;
        for (uint32_t body = 0; body < n; ++body) {
            const bpr::RigidWall &wall = ws[body];
            if (!(wall.geometry[0] != 0.0f || wall.geometry[1] != 0.0f)) continue;
            auto it = position.find((uint64_t)wall.owner);
            if (it == position.end()) continue;
            const uint32_t other = it->second;
            const bool lower = body < other;
            const uint64_t first = lower ? bs[body].id : (uint64_t)wall.owner;
            const uint64_t second = lower ? (uint64_t)wall.owner : bs[body].id;
            if (!emitted.insert({first, second}).second) continue;
            bp_contact r{};
            r.first = (int64_t)first;
            r.second = (int64_t)second;
            r.kind = BP_CONTACT_RIGID;
            r.normal[0] = lower ? -wall.geometry[0] : wall.geometry[0];
            r.normal[1] = lower ? -wall.geometry[1] : wall.geometry[1];
            r.point[0] = wall.geometry[2];
            r.point[1] = wall.geometry[3];
            r.impulse = wall.extra[0];
            r.index_b = lower ? other : body;
            push(r);
        }
        *outCount = count;
        return BP_OK;
    }
};

static bp_status rigid_world_from_desc(HostRigidWorld &w, const bp_rigid_world_desc *d) {
    if (!d) return BP_ERR_INVALID_ARGUMENT;
    if ((d->body_count && !d->bodies) || (d->vertex_count && !d->vertices) || (d->triangle_count && !d->triangles) ||
        (d->segment_count && !d->segments) || (d->joint_count && !d->joints) ||
        (d->joint_reference_count && !d->joint_references) || (d->joint_partner_count && !d->joint_partners)) {
        set_error("a non-empty array is null");
        return BP_ERR_INVALID_ARGUMENT;
    }
    w.config = d->config;
    w.bodies.assign(d->bodies, d->bodies + d->body_count);
    w.vertices.assign(d->vertices, d->vertices + 2 * (size_t)d->vertex_count);
    w.triangles.assign(d->triangles, d->triangles + 4 * (size_t)d->triangle_count);
    w.segments.assign(d->segments, d->segments + d->segment_count);
    w.joints.assign(d->joints, d->joints + d->joint_count);
    w.references.assign(d->joint_references, d->joint_references + 2 * (size_t)d->joint_reference_count);
    if (d->joint_ranges) w.ranges.assign(d->joint_ranges, d->joint_ranges + 4 * (size_t)d->body_count);
    else w.ranges.clear();
    w.partners.assign(d->joint_partners, d->joint_partners + d->joint_partner_count);
    w.failed = false;
    w.lastOk = false;
    w.stepped = false;
    return validate(w);
}

extern "C" {

bp_status bp_rigid_batch_create(bp_rigid_batch **out) {
    if (!out) return BP_ERR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device is available");
        return BP_ERR_RESOURCE_UNAVAILABLE;
    }
    *out = new bp_rigid_batch();
    return BP_OK;
}

void bp_rigid_batch_destroy(bp_rigid_batch *b) { delete b; }

bp_status bp_rigid_batch_add_world(bp_rigid_batch *b, const bp_rigid_world_desc *desc, uint32_t *out_index) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    HostRigidWorld w;
    bp_status s = rigid_world_from_desc(w, desc);
    if (s != BP_OK) return s;
    s = b->sync_from_device();
    if (s != BP_OK) return s;
    b->worlds.push_back(std::move(w));
    b->packed = false;
    if (out_index) *out_index = (uint32_t)(b->worlds.size() - 1);
    return BP_OK;
}

uint32_t bp_rigid_batch_world_count(const bp_rigid_batch *b) { return b ? (uint32_t)b->worlds.size() : 0u; }

bp_status bp_rigid_batch_set_world(bp_rigid_batch *b, uint32_t world, const bp_rigid_world_desc *desc) {
    if (!b || world >= b->worlds.size()) return BP_ERR_INVALID_ARGUMENT;
    HostRigidWorld w;
    bp_status s = rigid_world_from_desc(w, desc);
    if (s != BP_OK) return s;
    s = b->sync_from_device();
    if (s != BP_OK) return s;
    b->worlds[world] = std::move(w);
    b->packed = false;
    return BP_OK;
}

bp_status bp_rigid_batch_step(bp_rigid_batch *b, uint32_t steps, bp_rigid_frame_info *infos) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    return b->step(steps, infos);
}

bp_status bp_rigid_batch_read_bodies(bp_rigid_batch *b, uint32_t world, bp_rigid_body *out, uint32_t count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    return b->read_bodies(world, out, count);
}

bp_status bp_rigid_batch_read_joints(bp_rigid_batch *b, uint32_t world, bp_joint *out, uint32_t count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    return b->read_joints(world, out, count);
}

bp_status bp_rigid_batch_contacts(bp_rigid_batch *b, uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    return b->contacts(world, out, capacity, out_count);
}

float bp_rigid_batch_last_gpu_milliseconds(const bp_rigid_batch *b) { return b ? b->lastMs : 0.0f; }

// NativeJointState.prepare (NativeJointState.swift:73-141) for a world without bubbles: the
// entity of a body is its index.
bp_status bp_rigid_joint_tables(const bp_joint_desc *joints, uint32_t joint_count, const uint64_t *body_ids,
                                uint32_t body_count, bp_joint *records, uint32_t *out_record_count, uint32_t *references,
                                uint32_t *out_reference_count, uint32_t *ranges, uint64_t *partners,
                                uint32_t *out_partner_count) {
    if ((joint_count && !joints) || (body_count && (!body_ids || !ranges)) || !out_record_count || !out_reference_count ||
        !out_partner_count || (joint_count && (!records || !references || !partners)))
        return BP_ERR_INVALID_ARGUMENT;
    std::map<uint64_t, uint32_t> entity;
    for (uint32_t i = 0; i < body_count; ++i) entity[body_ids[i]] = i;
    std::vector<const bp_joint_desc *> sorted;
    for (uint32_t j = 0; j < joint_count; ++j) sorted.push_back(&joints[j]);
    std::stable_sort(sorted.begin(), sorted.end(), [](const bp_joint_desc *a, const bp_joint_desc *b) { return a->id < b->id; });
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> lists(body_count);
    std::vector<std::vector<uint64_t>> partnerLists(body_count);
    uint32_t count = 0;
    for (const bp_joint_desc *j : sorted) {
        auto first = entity.find(j->body_a);
        if (first == entity.end()) continue;
        uint32_t endA = first->second, endB = BP_JOINT_WORLD_END;
        if (j->has_body_b) {
            auto second = entity.find(j->body_b);
            if (second == entity.end()) continue;
            endB = second->second;
        }
        if (j->kind > BP_JOINT_WELD) { set_error("unknown joint kind"); return BP_ERR_INVALID_SNAPSHOT; }
        float anchors[4] = {j->anchor_a[0], j->anchor_a[1], j->anchor_b[0], j->anchor_b[1]};
        uint32_t e0 = endA, e1 = endB;
        if (!j->has_body_b) {
            // A joint to the world: the world is end 0 and the body end 1.
            e0 = BP_JOINT_WORLD_END;
            e1 = endA;
            anchors[0] = j->anchor_b[0];
            anchors[1] = j->anchor_b[1];
            anchors[2] = j->anchor_a[0];
            anchors[3] = j->anchor_a[1];
        }
        bp_joint r{};
        r.id = j->id;
        r.ends[0] = e0;
        r.ends[1] = e1;
        r.ends[2] = j->kind;
        std::memcpy(r.anchors, anchors, sizeof(anchors));
        if (j->kind == BP_JOINT_DISTANCE) {
            r.parameters[0] = j->length;
            r.parameters[1] = j->softness;
            r.parameters[2] = j->damping;
        } else {
            r.parameters[3] = j->reference_angle;
            if (j->kind == BP_JOINT_HINGE) {
                if (j->flags & BP_JOINT_HAS_LIMITS) {
                    r.ends[3] |= 1u;
                    r.limits[0] = j->lower;
                    r.limits[1] = j->upper;
                }
                if (j->flags & BP_JOINT_HAS_MOTOR) {
                    r.ends[3] |= 2u;
                    r.limits[2] = j->motor_speed;
                    r.limits[3] = j->max_torque;
                }
            }
        }
        const uint32_t index = count++;
        records[index] = r;
        if (e0 != BP_JOINT_WORLD_END) lists[e0].push_back({index, 0u});
        lists[e1].push_back({index, 1u});
        if (!j->collide_connected && j->has_body_b) {
            partnerLists[endA].push_back(j->body_b);
            partnerLists[endB].push_back(j->body_a);
        }
    }
    uint32_t refCursor = 0, partnerCursor = 0;
    for (uint32_t i = 0; i < body_count; ++i) {
        std::set<uint64_t> distinct(partnerLists[i].begin(), partnerLists[i].end());
        ranges[4 * i + 0] = refCursor;
        ranges[4 * i + 1] = (uint32_t)lists[i].size();
        ranges[4 * i + 2] = partnerCursor;
        ranges[4 * i + 3] = (uint32_t)distinct.size();
        for (auto &ref : lists[i]) {
            references[2 * refCursor] = ref.first;
            references[2 * refCursor + 1] = ref.second;
            ++refCursor;
        }
        for (uint64_t p : distinct) partners[partnerCursor++] = p;
    }
    *out_record_count = count;
    *out_reference_count = refCursor;
    *out_partner_count = partnerCursor;
    return BP_OK;
}

}  // extern "C"
