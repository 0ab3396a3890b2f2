// bp_api.cpp -- host side of the C ABI (include/bubblephysics.h).
//
// A bp_batch owns one device arena holding every world in structure-of-arrays form. Worlds
// are grouped by size class; each class (workgroup width 64, 256, 512 or 1024) runs on its
// own HIP stream, because worlds never exchange data: a class advances all its steps without
// waiting for the others. A bp_world is a batch of one world.
//
// Publication follows NativeMetalSolver.executeSnapshot (NativeMetalSolver.swift:1179-1201):
// the kernel writes the step into the unpublished side of the state and flips the published
// side only when the step succeeded, so a failed step leaves the last published state intact.
#include "bp_batch.h"

using namespace bph;


namespace {

bp_status validate_world(const HostWorld &w) {
    const bp_world_config &c = w.config;
    if (!(c.time_step > 0.0f) || !std::isfinite(c.time_step)) { set_error("time_step must be finite and positive"); return BP_ERR_INVALID_SNAPSHOT; }
    for (size_t i = 1; i < w.bubbles.size(); ++i)
        if (!(w.bubbles[i - 1].id < w.bubbles[i].id)) { set_error("bubbles must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
    for (size_t i = 1; i < w.segments.size(); ++i)
        if (!(w.segments[i - 1].id < w.segments[i].id)) { set_error("segments must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
    for (const bp_body &b : w.bodies) {
        if ((b.flags & (BP_BODY_STATIC | BP_BODY_KINEMATIC)) == 0 && !w.bubbles.empty()) {
            set_error("dynamic rigid bodies cannot share a world with bubbles");
            return BP_ERR_INVALID_SNAPSHOT;
        }
        if ((uint64_t)b.vertex_offset + b.vertex_count > w.vertices.size() / 2 ||
            (uint64_t)b.triangle_offset + b.triangle_count > w.triangles.size() / 4) {
            set_error("body geometry range outside the vertex or triangle arrays");
            return BP_ERR_INVALID_SNAPSHOT;
        }
        if (b.vertex_count == 0) { set_error("body without vertices"); return BP_ERR_INVALID_SNAPSHOT; }
        // The bounding radius holds the body (its vertices, with the core radius around a round
        // body's core): the kernels skip a body's exact tests for centres outside that circle.
        double extent = 0.0;
        const uint32_t corners = b.shape == BP_SHAPE_POLYGON ? b.vertex_count : std::min<uint32_t>(b.vertex_count, 2);
        for (uint32_t v = 0; v < corners; ++v) {
            const double x = w.vertices[2 * ((size_t)b.vertex_offset + v)], y = w.vertices[2 * ((size_t)b.vertex_offset + v) + 1];
            extent = std::max(extent, std::sqrt(x * x + y * y));
        }
        if (b.shape != BP_SHAPE_POLYGON) extent += (double)b.core_radius;
        if (!((double)b.bounding_radius >= extent * (1.0 - 1e-6) - 1e-6)) {
            set_error("body " + std::to_string(b.id) + ": bounding_radius " + std::to_string(b.bounding_radius) +
                      " is smaller than the body's extent " + std::to_string(extent));
            return BP_ERR_INVALID_SNAPSHOT;
        }
    }
    for (const bp_body &b : w.bodies)
        for (uint32_t t = 0; t < b.triangle_count; ++t)
            for (int k = 0; k < 3; ++k)
                if (w.triangles[4 * ((size_t)b.triangle_offset + t) + k] >= b.vertex_count) {
                    set_error("triangle index outside its body's vertices");
                    return BP_ERR_INVALID_SNAPSHOT;
                }
    for (const bp_field &f : w.fields) {
        if (f.kind > BP_FIELD_VORTEX) { set_error("unknown force field kind"); return BP_ERR_INVALID_SNAPSHOT; }
        if ((uint64_t)f.vertex_offset + f.vertex_count > w.fieldVertices.size() / 2 ||
            (f.has_filter && (uint64_t)f.filter_offset + f.filter_count > w.fieldIds.size())) {
            set_error("force field range outside the field vertex or filter arrays");
            return BP_ERR_INVALID_SNAPSHOT;
        }
        if (f.has_filter)
            for (uint32_t k = 1; k < f.filter_count; ++k)
                if (!((uint64_t)w.fieldIds[f.filter_offset + k - 1] < (uint64_t)w.fieldIds[f.filter_offset + k])) {
                    set_error("force field filter ids must be strictly increasing");
                    return BP_ERR_INVALID_SNAPSHOT;
                }
    }
    if (w.bubbles.size() > (1u << 30) || w.segments.size() > (1u << 30) || w.bodies.size() > (1u << 30)) {
        set_error("world too large"); return BP_ERR_CAPACITY_OVERFLOW;
    }
    return BP_OK;
}

}  // namespace

namespace {

// The controls of a world's descriptor from its configuration (offsets and counts are set by the
// caller).
void desc_controls(const HostWorld &w, uint32_t n, bpd::WorldDesc &d) {
    const bp_world_config &c = w.config;
    d.dt = c.time_step; d.contactTol = c.contact_tolerance; d.sepTol = c.separation_tolerance;
    for (int k = 0; k < 4; ++k) d.region[k] = c.region[k];
    d.globalForce[0] = c.global_force[0]; d.globalForce[1] = c.global_force[1];
    d.applyRegion = c.apply_region;
    d.limit = std::min(std::max(c.candidate_limit, 1u), 128u);
    d.split = std::max(c.substep_split, 1u);
    d.minimumMarginOnly = c.minimum_margin_only;
    d.buckets = c.grid_bucket_count == 0 ? auto_buckets(n) : std::min(c.grid_bucket_count, bpd::kMaxBuckets);
    d.assertEverySubstep = c.assert_every_substep;
    d.injectionEnabled = c.injection_enabled; d.injectionAfter = c.injection_after_substep;
    d.injectionObject = c.injection_object;
    d.injectionPosition[0] = c.injection_position[0]; d.injectionPosition[1] = c.injection_position[1];
    d.failed = w.failed ? 1u : 0u;
    d.exportAdjacency = (w.exportAdjacency || w.config.export_adjacency) ? 1u : 0u;
}

// A world's geometry and force records in their device form (segments, bodies, vertices,
// triangles, motors and bands sorted by bubble id, fields sorted by id).
struct GeometryStage {
    std::vector<float4> segAB, segRegion, bodyPose, bodyMotion, bodyShape, motors, bands;
    std::vector<uint32_t> segFlags, segCat, segMask;
    std::vector<int64_t> segId, segOwner;
    std::vector<float2> bodyPrev, bodyLimits, verts, fieldVerts;
    std::vector<uint4> bodyRanges, bodyKind, tris;
    std::vector<uint64_t> bodyId, fieldIds;
    std::vector<bpd::DevField> fields;
};

void stage_geometry(const HostWorld &w, GeometryStage &g) {
    const size_t ns = w.segments.size(), nb = w.bodies.size();
    g.segAB.resize(ns); g.segRegion.resize(ns); g.segFlags.resize(ns); g.segCat.resize(ns); g.segMask.resize(ns);
    g.segId.resize(ns); g.segOwner.resize(ns);
    for (size_t j = 0; j < ns; ++j) {
        const bp_segment &sg = w.segments[j];
        g.segAB[j] = make_float4(sg.a[0], sg.a[1], sg.b[0], sg.b[1]);
        g.segRegion[j] = make_float4(sg.previous_start[0], sg.previous_start[1], sg.outward_sign, sg.allowed_side);
        g.segFlags[j] = sg.flags; g.segCat[j] = sg.category; g.segMask[j] = sg.mask; g.segId[j] = sg.id; g.segOwner[j] = sg.owner_id;
    }
    g.bodyPose.resize(nb); g.bodyPrev.resize(nb); g.bodyMotion.resize(nb); g.bodyShape.resize(nb); g.bodyLimits.resize(nb);
    g.bodyRanges.resize(nb); g.bodyKind.resize(nb); g.bodyId.resize(nb);
    for (size_t k = 0; k < nb; ++k) {
        const bp_body &b = w.bodies[k];
        g.bodyPose[k] = make_float4(b.pose[0], b.pose[1], b.pose[2], b.previous[2]);
        g.bodyPrev[k] = make_float2(b.previous[0], b.previous[1]);
        g.bodyMotion[k] = make_float4(b.velocity[0], b.velocity[1], b.velocity[2], 0.0f);
        g.bodyShape[k] = make_float4(b.bounding_radius, b.core_radius, b.half_length, 0.0f);
        g.bodyLimits[k] = make_float2(b.max_linear_speed, b.max_angular_speed);
        g.bodyRanges[k] = make_uint4(b.vertex_offset, b.vertex_count, b.triangle_offset, b.triangle_count);
        g.bodyKind[k] = make_uint4(b.shape, b.flags, b.category, b.mask);
        g.bodyId[k] = b.id;
    }
    g.verts.resize(w.vertices.size() / 2);
    for (size_t v = 0; v < g.verts.size(); ++v) g.verts[v] = make_float2(w.vertices[2 * v], w.vertices[2 * v + 1]);
    g.tris.resize(w.triangles.size() / 4);
    for (size_t t = 0; t < g.tris.size(); ++t)
        g.tris[t] = make_uint4(w.triangles[4 * t], w.triangles[4 * t + 1], w.triangles[4 * t + 2], w.triangles[4 * t + 3]);
    std::vector<bp_motor> ms = w.motors;
    std::sort(ms.begin(), ms.end(), [](const bp_motor &a, const bp_motor &b) { return a.bubble_id < b.bubble_id; });
    g.motors.resize(ms.size());
    for (size_t k = 0; k < ms.size(); ++k) g.motors[k] = make_float4(ms[k].force[0], ms[k].force[1], ms[k].torque, ms[k].inertia);
    std::vector<bp_band> bs = w.bands;
    std::sort(bs.begin(), bs.end(), [](const bp_band &a, const bp_band &b) { return a.bubble_id < b.bubble_id; });
    g.bands.resize(bs.size());
    for (size_t k = 0; k < bs.size(); ++k)
        g.bands[k] = make_float4(bs[k].target[0], bs[k].target[1], bs[k].stiffness, bs[k].maximum_force);
    // Fields act in ascending id order (NativeForceInput sorts its objects, NativeForceRecords.swift:187).
    std::vector<bp_field> fs = w.fields;
    std::stable_sort(fs.begin(), fs.end(), [](const bp_field &a, const bp_field &b) { return a.id < b.id; });
    g.fields.resize(fs.size());
    for (size_t k = 0; k < fs.size(); ++k) {
        const bp_field &f = fs[k];
        g.fields[k].pose = make_float4(f.position[0], f.position[1], f.rotation, f.falloff);
        g.fields[k].strength = make_float4(f.direction[0], f.direction[1], f.strength, f.radius);
        g.fields[k].shape = make_uint4(f.kind, f.vertex_offset, f.vertex_count, 0u);
        g.fields[k].filter = make_uint4(f.filter_offset, f.filter_count, f.has_filter ? 1u : 0u, 0u);
    }
    g.fieldVerts.resize(w.fieldVertices.size() / 2);
    for (size_t v = 0; v < g.fieldVerts.size(); ++v) g.fieldVerts[v] = make_float2(w.fieldVertices[2 * v], w.fieldVertices[2 * v + 1]);
    g.fieldIds.resize(w.fieldIds.size());
    for (size_t k = 0; k < w.fieldIds.size(); ++k) g.fieldIds[k] = (uint64_t)w.fieldIds[k];
}

template <class T> void place(std::vector<T> &dst, size_t offset, const std::vector<T> &src) {
    std::copy(src.begin(), src.end(), dst.begin() + offset);
}

}  // namespace

bp_status bp_batch::pack() {
    // A failed world stays failed across a repack (it is not advanced again until it receives
    // new input): keep the device's sticky flags of the worlds packed last time.
    if (descs && !hostDescs.empty()) {
        std::vector<bpd::WorldDesc> cur(std::min(hostDescs.size(), worlds.size()));
        if (!cur.empty() && hipMemcpy(cur.data(), descs, cur.size() * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost) == hipSuccess)
            for (size_t w = 0; w < cur.size(); ++w) {
                worlds[w].failed = cur[w].failed == bpd::kWorldFailed;
                if (cur[w].failed == bpd::kWorldHeld) {
                    // A held world's removed bubbles are still in its published range; the repack
                    // reads the state back first (download_state), so delete them from the mirror.
                    std::vector<uint32_t> region(cur[w].bubbleCount);
                    std::vector<int64_t> ids(cur[w].bubbleCount);
                    if (!region.empty() &&
                        hipMemcpy(region.data(), arena.region[cur[w].pub] + cur[w].bubbleOffset, region.size() * 4,
                                  hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(ids.data(), arena.id + cur[w].bubbleOffset, ids.size() * 8, hipMemcpyDeviceToHost) == hipSuccess) {
                        auto &bs = worlds[w].bubbles;
                        for (size_t i = 0; i < region.size(); ++i)
                            if (region[i] == BP_REGION_REMOVED) {
                                const int64_t id = ids[i];
                                bs.erase(std::remove_if(bs.begin(), bs.end(), [&](const bp_bubble &b) { return b.id == id; }), bs.end());
                            }
                    }
                }
            }
    }
    release();
    uint32_t nb = 0, ns = 0, nbod = 0, nv = 0, nt = 0, nm = 0, nband = 0, nf = 0, nfv = 0, nfi = 0;
    uint64_t adj = 0, coopCells = 0;
    uint32_t coopWorlds = 0;
    const uint32_t coopMax = bp_coop_max_groups();
    std::vector<int> worldClass(worlds.size());
    // Geometry ranges: each world's ranges hold the batch's largest counts up to a small bound, so
    // that a served world can take over the ranges of an evicted one without a repack.
    uint32_t gSeg = 0, gBody = 0, gVert = 0, gTri = 0;
    for (const HostWorld &w : worlds) {
        gSeg = std::max(gSeg, (uint32_t)std::min<size_t>(w.segments.size(), kReserveSegments));
        gBody = std::max(gBody, (uint32_t)std::min<size_t>(w.bodies.size(), kReserveBodies));
        gVert = std::max(gVert, (uint32_t)std::min<size_t>(w.vertices.size() / 2, kReserveVertices));
        gTri = std::max(gTri, (uint32_t)std::min<size_t>(w.triangles.size() / 4, kReserveTriangles));
    }
    auto segRange = [&](const HostWorld &w) { return std::max((uint32_t)w.segments.size(), gSeg); };
    auto bodyRange = [&](const HostWorld &w) { return std::max((uint32_t)w.bodies.size(), gBody); };
    auto vertRange = [&](const HostWorld &w) { return std::max((uint32_t)(w.vertices.size() / 2), gVert); };
    auto triRange = [&](const HostWorld &w) { return std::max((uint32_t)(w.triangles.size() / 4), gTri); };
    coopGroups.assign(worlds.size(), 0);
    for (size_t wi = 0; wi < worlds.size(); ++wi) {
        const HostWorld &w = worlds[wi];
        bp_status s = validate_world(w);
        if (s != BP_OK) return s;
        int cls = default_class_for(w, coopMax > 0);
        if (w.forcedClass >= 0 && class_fits(w.forcedClass, w) && (w.forcedClass != kCoopClass || coopMax > 0)) cls = w.forcedClass;
        if (!w.resident) cls = -1;
        worldClass[wi] = cls;
        // Slots for the world's bubbles, with room to grow in place (serving transactions); a world
        // that outgrows its kernel configuration moves to a wider one without moving its data.
        const uint32_t n0 = (uint32_t)w.bubbles.size();
        const uint32_t slots = w.resident ? n0 + std::max<uint32_t>(8, n0 / 8) : 0u;
        worlds[wi].capacity = slots;
        worlds[wi].packedClass = cls;
        if (cls == kCoopClass) {
            coopCells += next_pow2((uint32_t)w.bubbles.size(), 1024);
            ++coopWorlds;
            const uint32_t block = bp_coop_block();
            coopGroups[wi] = std::max<uint32_t>(1, std::min<uint32_t>(coopMax, ((uint32_t)w.bubbles.size() + block - 1) / block));
        }
        nb += slots; ns += segRange(w); nbod += bodyRange(w);
        nv += vertRange(w); nt += triRange(w);
        nm += (uint32_t)w.motors.size(); nband += (uint32_t)w.bands.size();
        nf += (uint32_t)w.fields.size(); nfv += (uint32_t)(w.fieldVertices.size() / 2); nfi += (uint32_t)w.fieldIds.size();
        adj += w.resident ? (uint64_t)std::max<uint32_t>(slots, 16) * w.adjPerBubble : 0;
    }
    if (adj > 0xffffffffull) { set_error("adjacency capacity exceeds 2^32 entries"); return BP_ERR_CAPACITY_OVERFLOW; }
    totalBubbles = nb;
    totalBodies = nbod;
    bpd::Arena &A = arena;
    bp_status s = BP_OK;
#define BP_ALLOC(field, count) do { if ((s = alloc(field, count)) != BP_OK) return s; } while (0)
    for (int side = 0; side < 2; ++side) {
        BP_ALLOC(A.pos[side], nb); BP_ALLOC(A.start[side], nb); BP_ALLOC(A.vel[side], nb);
        BP_ALLOC(A.rot[side], nb); BP_ALLOC(A.omega[side], nb); BP_ALLOC(A.region[side], nb);
        BP_ALLOC(A.bodyPose[side], nbod); BP_ALLOC(A.bodyPrev[side], nbod);
    }
    BP_ALLOC(A.id, nb); BP_ALLOC(A.mass, nb); BP_ALLOC(A.invm, nb); BP_ALLOC(A.radius, nb); BP_ALLOC(A.Kc, nb);
    BP_ALLOC(A.gamma, nb); BP_ALLOC(A.flags, nb); BP_ALLOC(A.cat, nb); BP_ALLOC(A.mask, nb);
    BP_ALLOC(A.motor, nb); BP_ALLOC(A.band, nb);
    BP_ALLOC(A.acc, nb); BP_ALLOC(A.a0, nb); BP_ALLOC(A.ref, nb); BP_ALLOC(A.margin, nb);
    BP_ALLOC(A.rowStart, nb); BP_ALLOC(A.rowLen, nb); BP_ALLOC(A.grid, nb);
    BP_HIP(hipMemset(A.rowStart, 0, std::max<size_t>(nb, 1) * sizeof(uint32_t)));
    BP_HIP(hipMemset(A.rowLen, 0, std::max<size_t>(nb, 1) * sizeof(uint32_t)));
    BP_ALLOC(A.omega2, 1); BP_ALLOC(A.bound, 1); BP_ALLOC(A.bodyHits, nbod); BP_ALLOC(A.bodyCut, nbod);
    BP_ALLOC(A.nbrA, nb); BP_ALLOC(A.nbrB, nb); BP_ALLOC(A.invmEff, nb);
    BP_ALLOC(A.cellA, nb); BP_ALLOC(A.cellB, nb);
    BP_ALLOC(A.adj, adj);
    BP_ALLOC(A.adjGeo, adj);
    BP_ALLOC(A.segAB, ns); BP_ALLOC(A.segRegion, ns); BP_ALLOC(A.segFlags, ns); BP_ALLOC(A.segCat, ns);
    BP_ALLOC(A.segMask, ns); BP_ALLOC(A.segId, ns); BP_ALLOC(A.segOwner, ns);
    BP_ALLOC(A.bodyMotion, nbod); BP_ALLOC(A.bodyShape, nbod); BP_ALLOC(A.bodyLimits, nbod);
    BP_ALLOC(A.bodyRanges, nbod); BP_ALLOC(A.bodyKind, nbod); BP_ALLOC(A.bodyId, nbod);
    BP_ALLOC(A.verts, nv); BP_ALLOC(A.tris, nt);
    BP_ALLOC(A.motors, nm); BP_ALLOC(A.bands, nband);
    BP_ALLOC(A.fields, nf); BP_ALLOC(A.fieldVerts, nfv); BP_ALLOC(A.fieldIds, nfi);
    BP_ALLOC(A.history, (size_t)worlds.size() * bpd::kFrameLimit);
    BP_ALLOC(A.coopCells, coopCells); BP_ALLOC(A.coopScratch, (size_t)coopWorlds * bpd::kCoopScratchWords);
    BP_ALLOC(A.bvhChild, 2 * (size_t)nbod); BP_ALLOC(A.bvhBox, 2 * (size_t)nbod); BP_ALLOC(A.bvhLeaf, nbod);
    BP_ALLOC(descs, worlds.size());
    BP_ALLOC(arenaDev, 1);
    for (int k = 0; k < kClassCount; ++k) BP_ALLOC(orders[k], worlds.size());
    BP_ALLOC(subsetOrders, worlds.size());
    BP_ALLOC(cursors, kClassCount + 1);
    xferCapacity = 1;
    for (const auto &w : worlds) xferCapacity = std::max(xferCapacity, w.capacity);
    BP_ALLOC(xferBubbles, xferCapacity);
    BP_ALLOC(xferIndices, 2 * (size_t)xferCapacity);
#undef BP_ALLOC

    // Host staging of every array.
    std::vector<float2> pos(nb), start(nb), vel(nb), bpose2(nbod), verts(nv), zero2(nb, make_float2(0, 0));
    std::vector<float> rot(nb), om(nb), mass(nb), invm(nb), radius(nb), Kc(nb), gamma(nb);
    std::vector<uint32_t> flags(nb), cat(nb), mask(nb), region(nb, 0);
    std::vector<int32_t> motorIdx(nb, -1), bandIdx(nb, -1);
    std::vector<int64_t> ids(nb), segId(ns), segOwner(ns);
    std::vector<float4> segAB(ns), segRegion(ns), bodyPose(nbod), bodyMotion(nbod), bodyShape(nbod), motors(nm), bands(nband);
    std::vector<uint32_t> segFlags(ns), segCat(ns), segMask(ns);
    std::vector<float2> bodyLimits(nbod);
    std::vector<uint4> bodyRanges(nbod), bodyKind(nbod), tris(nt);
    std::vector<bpd::DevField> fields(nf);
    std::vector<float2> fieldVerts(nfv);
    std::vector<uint64_t> fieldIds(nfi);
    std::vector<uint64_t> bodyId(nbod);
    hostDescs.assign(worlds.size(), bpd::WorldDesc{});
    regions.assign(worlds.size(), Region{});
    uint32_t ob = 0, os = 0, obd = 0, ov = 0, ot = 0, om_ = 0, oband = 0, of = 0, ofv = 0, ofi = 0, occ = 0, ocs = 0;
    uint64_t oadj = 0;
    for (auto &k : classWorlds) k.clear();
    for (size_t wi = 0; wi < worlds.size(); ++wi) {
        const HostWorld &w = worlds[wi];
        bpd::WorldDesc &d = hostDescs[wi];
        uint32_t n = (uint32_t)w.bubbles.size();
        d.bubbleOffset = ob; d.bubbleCount = w.resident ? n : 0;
        d.segmentOffset = os; d.segmentCount = (uint32_t)w.segments.size();
        d.bodyOffset = obd; d.bodyCount = (uint32_t)w.bodies.size();
        d.vertexOffset = ov; d.triangleOffset = ot;
        d.vertexCount = (uint32_t)(w.vertices.size() / 2); d.triangleCount = (uint32_t)(w.triangles.size() / 4);
        d.motorOffset = om_; d.motorCount = (uint32_t)w.motors.size();
        d.bandOffset = oband; d.bandCount = (uint32_t)w.bands.size();
        d.fieldOffset = of; d.fieldCount = (uint32_t)w.fields.size(); d.fieldVertexOffset = ofv; d.fieldIdOffset = ofi;
        d.adjOffset = (uint32_t)oadj; d.adjCapacity = w.resident ? std::max<uint32_t>(w.capacity, 16) * w.adjPerBubble : 0;
        d.historyOffset = (uint32_t)(wi * bpd::kFrameLimit);
        d.pub = 0;
        desc_controls(w, n, d);
        d.bubbleCapacity = w.capacity;
        const uint32_t resident = w.resident ? n : 0u;       // bubbles staged in the device range
        for (uint32_t i = 0; i < resident; ++i) {
            const bp_bubble &b = w.bubbles[i];
            uint32_t g = ob + i;
            pos[g] = make_float2(b.center[0], b.center[1]);
            start[g] = make_float2(b.previous_center[0], b.previous_center[1]);
            vel[g] = make_float2(b.velocity[0], b.velocity[1]);
            rot[g] = b.rotation; om[g] = b.angular_velocity;
            mass[g] = b.mass; invm[g] = b.inverse_mass; radius[g] = b.radius; Kc[g] = b.contact_stiffness;
            gamma[g] = b.damping; flags[g] = b.flags; cat[g] = b.category; mask[g] = b.mask; ids[g] = b.id;
        }
        // Motor and band records are resolved to bubble indices once (ids are unique and sorted).
        GeometryStage gs;
        stage_geometry(w, gs);
        const std::vector<bp_motor> &msRaw = w.motors;
        std::vector<int64_t> motorIds(msRaw.size()), bandIds(w.bands.size());
        for (size_t k = 0; k < msRaw.size(); ++k) motorIds[k] = msRaw[k].bubble_id;
        for (size_t k = 0; k < w.bands.size(); ++k) bandIds[k] = w.bands[k].bubble_id;
        std::sort(motorIds.begin(), motorIds.end());
        std::sort(bandIds.begin(), bandIds.end());
        for (uint32_t k = 0; k < motorIds.size(); ++k) {
            auto it = std::lower_bound(w.bubbles.begin(), w.bubbles.end(), motorIds[k],
                                       [](const bp_bubble &b, int64_t id) { return b.id < id; });
            if (w.resident && it != w.bubbles.end() && it->id == motorIds[k]) motorIdx[ob + (it - w.bubbles.begin())] = (int32_t)k;
        }
        for (uint32_t k = 0; k < bandIds.size(); ++k) {
            auto it = std::lower_bound(w.bubbles.begin(), w.bubbles.end(), bandIds[k],
                                       [](const bp_bubble &b, int64_t id) { return b.id < id; });
            if (w.resident && it != w.bubbles.end() && it->id == bandIds[k]) bandIdx[ob + (it - w.bubbles.begin())] = (int32_t)k;
        }
        place(motors, om_, gs.motors); place(bands, oband, gs.bands);
        place(fields, of, gs.fields); place(fieldVerts, ofv, gs.fieldVerts); place(fieldIds, ofi, gs.fieldIds);
        place(segAB, os, gs.segAB); place(segRegion, os, gs.segRegion); place(segFlags, os, gs.segFlags);
        place(segCat, os, gs.segCat); place(segMask, os, gs.segMask); place(segId, os, gs.segId); place(segOwner, os, gs.segOwner);
        place(bodyPose, obd, gs.bodyPose); place(bpose2, obd, gs.bodyPrev); place(bodyMotion, obd, gs.bodyMotion);
        place(bodyShape, obd, gs.bodyShape); place(bodyLimits, obd, gs.bodyLimits); place(bodyRanges, obd, gs.bodyRanges);
        place(bodyKind, obd, gs.bodyKind); place(bodyId, obd, gs.bodyId);
        place(verts, ov, gs.verts); place(tris, ot, gs.tris);
        regions[wi] = Region{w.capacity, segRange(w), bodyRange(w), vertRange(w), triRange(w),
                             (uint32_t)gs.motors.size(), (uint32_t)gs.bands.size(),
                             (uint32_t)gs.fields.size(), (uint32_t)gs.fieldVerts.size(), (uint32_t)gs.fieldIds.size(),
                             d.adjCapacity};
        const int cls = worldClass[wi];
        if (cls >= 0) classWorlds[cls].push_back((uint32_t)wi);
        if (cls == kCoopClass) {
            d.coopBuckets = next_pow2(n, 1024);
            d.coopCellOffset = occ; occ += d.coopBuckets;
            d.coopScratchOffset = ocs; ocs += bpd::kCoopScratchWords;
        }
        ob += w.capacity; os += segRange(w); obd += bodyRange(w);
        ov += vertRange(w); ot += triRange(w);
        om_ += (uint32_t)w.motors.size(); oband += (uint32_t)w.bands.size();
        of += (uint32_t)w.fields.size(); ofv += (uint32_t)(w.fieldVertices.size() / 2); ofi += (uint32_t)w.fieldIds.size();
        oadj += (uint64_t)d.adjCapacity;
    }
    // Longest-processing-time order inside each class: the largest worlds start first (until a
    // step has measured each world's device time, see launch()).
    workEstimate.assign(worlds.size(), 0);
    for (size_t wi = 0; wi < worlds.size(); ++wi) workEstimate[wi] = hostDescs[wi].bubbleCount;
    for (int k = 0; k < kClassCount; ++k) classOrder[k] = kClassCount - 1 - k;
    orderDirty = false;
    for (auto &k : classWorlds)
        std::stable_sort(k.begin(), k.end(), [&](uint32_t a, uint32_t b) { return hostDescs[a].bubbleCount > hostDescs[b].bubbleCount; });
    if ((s = upload(A.pos[0], pos)) || (s = upload(A.start[0], start)) || (s = upload(A.vel[0], vel)) ||
        (s = upload(A.rot[0], rot)) || (s = upload(A.omega[0], om)) || (s = upload(A.region[0], region)) ||
        (s = upload(A.id, ids)) || (s = upload(A.mass, mass)) || (s = upload(A.invm, invm)) || (s = upload(A.radius, radius)) ||
        (s = upload(A.Kc, Kc)) || (s = upload(A.gamma, gamma)) || (s = upload(A.flags, flags)) || (s = upload(A.cat, cat)) ||
        (s = upload(A.mask, mask)) || (s = upload(A.motor, motorIdx)) || (s = upload(A.band, bandIdx)) ||
        (s = upload(A.acc, zero2)) || (s = upload(A.segAB, segAB)) || (s = upload(A.segRegion, segRegion)) ||
        (s = upload(A.segFlags, segFlags)) || (s = upload(A.segCat, segCat)) || (s = upload(A.segMask, segMask)) ||
        (s = upload(A.segId, segId)) || (s = upload(A.segOwner, segOwner)) || (s = upload(A.bodyPose[0], bodyPose)) ||
        (s = upload(A.bodyPrev[0], bpose2)) || (s = upload(A.bodyMotion, bodyMotion)) || (s = upload(A.bodyShape, bodyShape)) ||
        (s = upload(A.bodyLimits, bodyLimits)) || (s = upload(A.bodyRanges, bodyRanges)) || (s = upload(A.bodyKind, bodyKind)) ||
        (s = upload(A.bodyId, bodyId)) || (s = upload(A.verts, verts)) || (s = upload(A.tris, tris)) ||
        (s = upload(A.motors, motors)) || (s = upload(A.bands, bands)) || (s = upload(A.fields, fields)) ||
        (s = upload(A.fieldVerts, fieldVerts)) || (s = upload(A.fieldIds, fieldIds)) || (s = upload(descs, hostDescs)))
        return s;
    for (int k = 0; k < kClassCount; ++k)
        if ((s = upload(orders[k], classWorlds[k])) != BP_OK) return s;
    BP_HIP(hipMemcpy(arenaDev, &arena, sizeof arena, hipMemcpyHostToDevice));
    if ((s = create_streams()) != BP_OK) return s;
    if (!deviceCUs) {
        int device = 0;
        BP_HIP(hipGetDevice(&device));
        BP_HIP(hipDeviceGetAttribute(&deviceCUs, hipDeviceAttributeMultiprocessorCount, device));
        for (int k = 0; k < bp_lds_config_count(); ++k) residency[k] = bp_lds_config_residency(k);
    }
    for (auto &g : poolGroups) g = 0;
    estimatesMeasured = false;
    packed = true;
    deviceValid = true;
    return BP_OK;
}

// Workgroups of the pooled launches. The classes whose workgroup needs a whole CU (one resident
// per CU) share out the CUs in proportion to their measured work (the summed device time of
// their worlds' last steps), leaving the small classes a share in proportion to theirs, counted
// as if 1.5 of their workgroups ran on a CU at full speed (a small world's step is bound by
// latency, so company on a CU slows each one down; 1.5 is the measured best on config-4 batches).
// The small classes are launched one workgroup per world and fill the CUs the pools leave. A
// class whose share is at least its world count is launched one workgroup per world too.
void bp_batch::plan_pools() {
    for (auto &g : poolGroups) g = 0;
    if (!pooling || !estimatesMeasured || deviceCUs <= 0) return;
    double work[kClassCount] = {}, heavy = 0.0, small = 0.0;
    for (int k = 0; k < bp_lds_config_count(); ++k) {
        if (classWorlds[k].empty() || residency[k] <= 0) continue;
        for (uint32_t wi : classWorlds[k]) work[k] += (double)workEstimate[wi];
        if (residency[k] == 1) heavy += work[k];
        else small += work[k] / std::min((double)residency[k], 1.5);
    }
    if (!(heavy > 0.0)) return;
    const double cus = std::floor(deviceCUs * heavy / (heavy + small));
    for (int k = 0; k < bp_lds_config_count(); ++k) {
        const uint32_t count = (uint32_t)classWorlds[k].size();
        if (!count || residency[k] != 1) continue;
        const uint32_t g = (uint32_t)std::max(1.0, std::floor(cus * work[k] / heavy));
        poolGroups[k] = g < count ? g : 0;
    }
}

// Copies the published state of every world back into the host mirror.
bp_status bp_batch::download_state(bool full) {
    if (!packed || !deviceValid) return BP_OK;
    BP_HIP(hipMemcpy(hostDescs.data(), descs, hostDescs.size() * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost));
    uint32_t nb = totalBubbles;
    std::vector<int64_t> ids;
    std::vector<float> mass, invm, radius, Kc, gamma;
    std::vector<uint32_t> flags, cat, mask;
    if (full && nb) {
        ids.resize(nb); mass.resize(nb); invm.resize(nb); radius.resize(nb); Kc.resize(nb); gamma.resize(nb);
        flags.resize(nb); cat.resize(nb); mask.resize(nb);
        BP_HIP(hipMemcpy(ids.data(), arena.id, nb * sizeof(int64_t), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(mass.data(), arena.mass, nb * sizeof(float), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(invm.data(), arena.invm, nb * sizeof(float), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(radius.data(), arena.radius, nb * sizeof(float), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(Kc.data(), arena.Kc, nb * sizeof(float), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(gamma.data(), arena.gamma, nb * sizeof(float), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(flags.data(), arena.flags, nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(cat.data(), arena.cat, nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(mask.data(), arena.mask, nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    std::vector<float2> pos[2], start[2], vel[2];
    std::vector<float> rot[2], om[2];
    const size_t nbod = totalBodies;
    std::vector<float4> bodyPose[2];
    std::vector<float2> bodyPrev[2];
    std::vector<float4> bodyMotion(nbod);
    for (int side = 0; side < 2; ++side) {
        pos[side].resize(nb); start[side].resize(nb); vel[side].resize(nb); rot[side].resize(nb); om[side].resize(nb);
        bodyPose[side].resize(nbod); bodyPrev[side].resize(nbod);
        if (nb) {
            BP_HIP(hipMemcpy(pos[side].data(), arena.pos[side], nb * sizeof(float2), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(start[side].data(), arena.start[side], nb * sizeof(float2), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(vel[side].data(), arena.vel[side], nb * sizeof(float2), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(rot[side].data(), arena.rot[side], nb * sizeof(float), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(om[side].data(), arena.omega[side], nb * sizeof(float), hipMemcpyDeviceToHost));
        }
        if (nbod) {
            BP_HIP(hipMemcpy(bodyPose[side].data(), arena.bodyPose[side], nbod * sizeof(float4), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(bodyPrev[side].data(), arena.bodyPrev[side], nbod * sizeof(float2), hipMemcpyDeviceToHost));
        }
    }
    if (nbod) BP_HIP(hipMemcpy(bodyMotion.data(), arena.bodyMotion, nbod * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t wi = 0; wi < worlds.size(); ++wi) {
        HostWorld &w = worlds[wi];
        if (!w.resident) continue;                         // its host copy is the state
        const bpd::WorldDesc &d = hostDescs[wi];
        uint32_t p = d.pub;
        if (full) w.bubbles.resize(std::min(d.bubbleCount, w.capacity));
        for (uint32_t i = 0; i < w.bubbles.size(); ++i) {
            uint32_t g = d.bubbleOffset + i;
            bp_bubble &b = w.bubbles[i];
            if (full) {
                b.id = ids[g]; b.mass = mass[g]; b.inverse_mass = invm[g]; b.radius = radius[g];
                b.contact_stiffness = Kc[g]; b.damping = gamma[g]; b.flags = flags[g]; b.category = cat[g]; b.mask = mask[g];
            }
            b.center[0] = pos[p][g].x; b.center[1] = pos[p][g].y;
            b.previous_center[0] = start[p][g].x; b.previous_center[1] = start[p][g].y;
            b.velocity[0] = vel[p][g].x; b.velocity[1] = vel[p][g].y;
            b.rotation = rot[p][g]; b.angular_velocity = om[p][g];
        }
        for (uint32_t k = 0; k < w.bodies.size(); ++k) {
            uint32_t g = d.bodyOffset + k;
            bp_body &b = w.bodies[k];
            b.pose[0] = bodyPose[p][g].x; b.pose[1] = bodyPose[p][g].y; b.pose[2] = bodyPose[p][g].z;
            b.previous[0] = bodyPrev[p][g].x; b.previous[1] = bodyPrev[p][g].y; b.previous[2] = bodyPose[p][g].w;
            b.velocity[0] = bodyMotion[g].x; b.velocity[1] = bodyMotion[g].y; b.velocity[2] = bodyMotion[g].z;
        }
    }
    return BP_OK;
}

static void fill_info(const bpd::WorldStatus &s, bp_frame_info *out) {
    std::memset(out, 0, sizeof *out);
    out->status = s.status;
    out->substep_count = s.substeps;
    out->grid_rebuild_count = s.rebuilds;
    out->candidate_pair_count = s.candidatePairs;
    out->candidate_overflow = s.candidateOverflow;
    out->max_candidates_per_object = s.maxCandidates;
    out->execution_path = BP_PATH_WORKGROUP;
    out->group_width = s.groupWidth;
    out->region_removals = s.removals;
    out->region_wraps = s.wraps;
    out->simulated_time = s.time[1];
    out->limiter_object = s.limiterObject;
    out->limiter_term = s.limiterTerm;
    out->adjacency_entries = s.adjacencyEntries;
    out->device_ticks = s.ticks;
    out->detail.object_ids[0] = s.violationObjects[0];
    out->detail.object_ids[1] = s.violationObjects[1];
    out->detail.violation_kind = (int32_t)s.violationKind;
    out->detail.edge_segment_id_low = s.violationKind == BP_VIOLATION_EDGE_SIDE ? (int32_t)(uint32_t)s.violationSegment : -1;
    out->detail.distance = s.violationKind == BP_VIOLATION_NONE || s.violationKind == BP_VIOLATION_INSIDE_BODY
                               ? std::nanf("") : s.violationDetail;
    out->detail.limiting_term = s.status == BP_ERR_SUBSTEP_LIMIT_EXCEEDED ? s.limiterTerm : BP_TERM_NONE;
    if (s.status == BP_ERR_SUBSTEP_LIMIT_EXCEEDED) {
        out->detail.object_ids[0] = s.limiterObject;
        out->detail.object_ids[1] = 0;
    }
}

// Copies the published bubble state of one resident world back into its host mirror (one
// gather kernel and one copy; body poses are refreshed by download_state before a repack).
bp_status bp_batch::download_world(uint32_t wi) {
    HostWorld &w = worlds[wi];
    if (!packed || !deviceValid || !w.resident) return BP_OK;
    const uint32_t n = (uint32_t)w.bubbles.size();
    if (n == 0) return BP_OK;
    std::vector<bp_bubble> rec(n);
    hipError_t e = bp_launch_gather_world(arenaDev, descs + wi, xferBubbles, n, streams[0]);
    if (e != hipSuccess) { set_error(std::string("gather: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipMemcpyAsync(rec.data(), xferBubbles, n * sizeof(bp_bubble), hipMemcpyDeviceToHost, streams[0]));
    BP_HIP(hipStreamSynchronize(streams[0]));
    for (uint32_t i = 0; i < n; ++i) {
        bp_bubble &b = w.bubbles[i];
        const bp_bubble &r = rec[i];
        b.center[0] = r.center[0]; b.center[1] = r.center[1];
        b.previous_center[0] = r.previous_center[0]; b.previous_center[1] = r.previous_center[1];
        b.velocity[0] = r.velocity[0]; b.velocity[1] = r.velocity[1];
        b.rotation = r.rotation; b.angular_velocity = r.angular_velocity;
    }
    return BP_OK;
}

// Replaces the bubbles of a world (a serving transaction). Within the world's slots and kernel
// configuration the device range is rewritten in place; otherwise the batch is repacked at the
// next step.
bp_status bp_batch::write_bubbles(uint32_t wi, const bp_bubble *bubbles, uint32_t count) {
    HostWorld &w = worlds[wi];
    for (uint32_t i = 1; i < count; ++i)
        if (!(bubbles[i - 1].id < bubbles[i].id)) { set_error("bubbles must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
    const bool inPlace = packed && w.resident && w.packedClass >= 0 && count <= w.capacity;
    if (!inPlace) {
        bp_status s = download_state();
        if (s != BP_OK) return s;
        w.bubbles.assign(bubbles, bubbles + count);
        // New input: the world may be stepped again (the repack keeps the other worlds' flags).
        w.failed = false;
        if (packed && descs && wi < hostDescs.size()) {
            bpd::WorldDesc d;
            BP_HIP(hipMemcpy(&d, descs + wi, sizeof d, hipMemcpyDeviceToHost));
            if (d.failed) { d.failed = 0; BP_HIP(hipMemcpy(descs + wi, &d, sizeof d, hipMemcpyHostToDevice)); }
        }
        packed = false;
        return BP_OK;
    }
    bpd::WorldDesc d = hostDescs[wi];
    BP_HIP(hipMemcpy(&d, descs + wi, sizeof d, hipMemcpyDeviceToHost));   // the device side holds `pub`
    std::vector<int32_t> indices(2 * (size_t)count, -1);
    motor_band_indices(wi, bubbles, count, indices.data());
    if (count) {
        BP_HIP(hipMemcpyAsync(xferBubbles, bubbles, count * sizeof(bp_bubble), hipMemcpyHostToDevice, streams[0]));
        BP_HIP(hipMemcpyAsync(xferIndices, indices.data(), indices.size() * sizeof(int32_t), hipMemcpyHostToDevice, streams[0]));
        hipError_t e = bp_launch_scatter_world(arenaDev, d.bubbleOffset, count, xferBubbles, xferIndices, xferIndices + count,
                                               streams[0]);
        if (e != hipSuccess) { set_error(std::string("scatter: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
        BP_HIP(hipStreamSynchronize(streams[0]));
    }
    written_desc(wi, count, d);
    BP_HIP(hipMemcpy(descs + wi, &d, sizeof d, hipMemcpyHostToDevice));
    return adopt_written(wi, bubbles, count, d);
}

// Motor and band indices of a world's new bubble list (indices[i] and indices[count + i], -1 for
// none): the records keep their bubble ids, and the device holds them sorted by bubble id, as
// pack() wrote them.
void bp_batch::motor_band_indices(uint32_t wi, const bp_bubble *bubbles, uint32_t count, int32_t *indices) const {
    const HostWorld &w = worlds[wi];
    for (size_t k = 0; k < 2 * (size_t)count; ++k) indices[k] = -1;
    std::vector<bp_motor> ms = w.motors;
    std::sort(ms.begin(), ms.end(), [](const bp_motor &a, const bp_motor &b) { return a.bubble_id < b.bubble_id; });
    std::vector<bp_band> bs = w.bands;
    std::sort(bs.begin(), bs.end(), [](const bp_band &a, const bp_band &b) { return a.bubble_id < b.bubble_id; });
    for (uint32_t i = 0; i < count; ++i) {
        const int64_t id = bubbles[i].id;
        auto m = std::lower_bound(ms.begin(), ms.end(), id, [](const bp_motor &r, int64_t v) { return r.bubble_id < v; });
        if (m != ms.end() && m->bubble_id == id) indices[i] = (int32_t)(m - ms.begin());
        auto b = std::lower_bound(bs.begin(), bs.end(), id, [](const bp_band &r, int64_t v) { return r.bubble_id < v; });
        if (b != bs.end() && b->bubble_id == id) indices[count + i] = (int32_t)(b - bs.begin());
    }
}

// The descriptor of a world whose list was rewritten in place with `count` bubbles.
void bp_batch::written_desc(uint32_t wi, uint32_t count, bpd::WorldDesc &d) const {
    const HostWorld &w = worlds[wi];
    d.bubbleCount = count;
    d.buckets = w.config.grid_bucket_count == 0 ? auto_buckets(count) : std::min(w.config.grid_bucket_count, bpd::kMaxBuckets);
    d.failed = 0;
}

// Host bookkeeping of a world whose list was rewritten in place on the device (its scatter and
// descriptor `d` are complete).
bp_status bp_batch::adopt_written(uint32_t wi, const bp_bubble *bubbles, uint32_t count, const bpd::WorldDesc &d) {
    HostWorld &w = worlds[wi];
    hostDescs[wi] = d;
    w.failed = false;
    w.bubbles.assign(bubbles, bubbles + count);
    // A world that no longer fits its kernel configuration moves to another one: only the order
    // lists of the two configurations change. A world without cooperative storage takes the
    // one-workgroup global kernel instead of the cooperative one.
    if (!class_fits(w.packedClass, w)) {
        int cls = default_class_for(w, d.coopBuckets != 0);
        if (w.forcedClass >= 0 && class_fits(w.forcedClass, w) && (w.forcedClass != kCoopClass || d.coopBuckets != 0)) cls = w.forcedClass;
        if (cls == kCoopClass && d.coopBuckets == 0) cls = kGlobalClass;
        const int old = w.packedClass;
        auto &from = classWorlds[old];
        from.erase(std::remove(from.begin(), from.end(), wi), from.end());
        auto &to = classWorlds[cls];
        to.insert(std::upper_bound(to.begin(), to.end(), wi, [&](uint32_t a, uint32_t b) {
                      return hostDescs[a].bubbleCount > hostDescs[b].bubbleCount;
                  }), wi);
        w.packedClass = cls;
        release_graph();                              // its launches were captured with the old class lists
        bp_status s2 = BP_OK;
        if ((s2 = upload(orders[old], from)) != BP_OK || (s2 = upload(orders[cls], to)) != BP_OK) return s2;
    }
    return BP_OK;
}

// Device bytes held by a world: its bubble slots in every per-bubble array, its adjacency
// capacity and its geometry.
uint64_t bp_batch::footprint(uint32_t wi) const {
    const HostWorld &w = worlds[wi];
    const uint32_t n = (uint32_t)w.bubbles.size();
    const uint64_t slots = (packed && w.resident) ? w.capacity : n + std::max<uint32_t>(8, n / 8);   // as pack() reserves
    const uint64_t perSlot = 2 * (3 * sizeof(float2) + 2 * sizeof(float) + sizeof(uint32_t))   // two-sided state
                             + sizeof(int64_t) + 6 * sizeof(float) + 3 * sizeof(uint32_t) + 2 * sizeof(int32_t)
                             + 3 * sizeof(float2) + sizeof(float) + 3 * sizeof(uint32_t);     // properties, scratch
    const uint64_t adjacency = std::max<uint64_t>(slots, 16) * w.adjPerBubble * (sizeof(uint32_t) + 4 * sizeof(float));
    const uint64_t geometry = w.segments.size() * 64 + w.bodies.size() * 160 + w.vertices.size() * sizeof(float) +
                              w.triangles.size() * sizeof(uint32_t) + (w.motors.size() + w.bands.size()) * sizeof(float4);
    return slots * perSlot + adjacency + geometry + bpd::kFrameLimit * sizeof(float) + sizeof(bpd::WorldDesc);
}

bool bp_batch::fits_ranges_of(uint32_t out, uint32_t in) const {
    if (!packed || out >= worlds.size() || in >= worlds.size() || out == in || regions.size() != worlds.size()) return false;
    const HostWorld &o = worlds[out], &w = worlds[in];
    // A world moved in gets no cooperative storage (it runs the one-workgroup kernels, as served
    // worlds do); one that asks for the cooperative kernel waits for a repack.
    if (!o.resident || w.resident || o.packedClass < 0 || w.config.force_cooperative) return false;
    const Region &r = regions[out];
    const uint32_t n = (uint32_t)w.bubbles.size();
    return n + std::max<uint32_t>(8, n / 8) <= r.bubbles &&                     // the room to grow pack() gives
           (uint64_t)std::max<uint32_t>(r.bubbles, 16) * w.adjPerBubble <= r.adjacency &&
           w.segments.size() <= r.segments && w.bodies.size() <= r.bodies && w.vertices.size() / 2 <= r.vertices &&
           w.triangles.size() / 4 <= r.triangles && w.motors.size() <= r.motors && w.bands.size() <= r.bands &&
           w.fields.size() <= r.fields && w.fieldVertices.size() / 2 <= r.fieldVertices && w.fieldIds.size() <= r.fieldIds;
}

// One exchange of a swap round: the resident world `out` leaves its ranges to the held world `in`.
struct bp_batch::SwapPair {
    uint32_t out, in, result;
    size_t readAt;                     // in the pinned read-back area
    uint32_t recAt;                    // in the transfer buffer and the pinned records
    MirrorRange mirror;
    bpd::WorldDesc dn, dOut;
    GeometryStage gs;
    size_t recsAt, idxAt;              // in the upload blob
};

// The pairs of the swap round in flight, and those of them whose upload is queued.
struct bp_batch::SwapRound {
    std::vector<SwapPair> pairs, upload;
};

bp_batch::SwapRound &bp_batch::swap_round() {
    if (!swapRound) swapRound = new SwapRound();
    return *swapRound;
}

namespace {
inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
// Read-back layout of one resident world: its descriptor, then its bodies' poses on both sides,
// their motion, and their previous poses on both sides.
inline size_t read_size(size_t nb) { return align16(sizeof(bpd::WorldDesc)) + nb * (3 * sizeof(float4) + 2 * sizeof(float2)); }
}  // namespace

bp_status bp_batch::ensure_pinned_blob(size_t bytes) {
    if (pinnedBlobBytes >= bytes) return BP_OK;
    if (pinnedBlob) (void)hipHostFree(pinnedBlob);
    pinnedBlob = nullptr; pinnedBlobDev = nullptr; pinnedBlobBytes = 0;
    BP_HIP(hipHostMalloc(&pinnedBlob, bytes, hipHostMallocDefault));
    BP_HIP(hipHostGetDevicePointer(&pinnedBlobDev, pinnedBlob, 0));
    pinnedBlobBytes = bytes;
    return BP_OK;
}

bp_status bp_batch::swap_begin(const uint32_t *outs, const uint32_t *ins, uint32_t count, const MirrorRange *mirrors,
                               bp_status *results, hipStream_t stream) {
    SwapRound &sr = swap_round();
    sr.pairs.clear();
    sr.upload.clear();
    if (pinnedCapacity < xferCapacity) {
        if (pinnedRecs) (void)hipHostFree(pinnedRecs);
        if (pinnedIdx) (void)hipHostFree(pinnedIdx);
        pinnedRecs = nullptr; pinnedIdx = nullptr; pinnedCapacity = 0;
        BP_HIP(hipHostMalloc((void **)&pinnedRecs, (size_t)xferCapacity * sizeof(bp_bubble), hipHostMallocDefault));
        BP_HIP(hipHostMalloc((void **)&pinnedIdx, 2 * (size_t)xferCapacity * sizeof(int32_t), hipHostMallocDefault));
        pinnedCapacity = xferCapacity;
    }
    // The pairs whose resident worlds' records fit the transfer buffer together; the others wait
    // for a later round.
    size_t readBytes = 0;
    uint32_t recs = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t out = outs[i], in = ins[i];
        results[i] = BP_ERR_CAPACITY_OVERFLOW;
        if (!fits_ranges_of(out, in) || regions[out].bubbles > xferCapacity) continue;
        const uint32_t rb = regions[out].bubbles;
        results[i] = BP_ERR_UNAVAILABLE;
        if (recs + rb > xferCapacity) continue;
        results[i] = BP_OK;
        SwapPair pr{};
        pr.out = out; pr.in = in; pr.result = i;
        pr.readAt = readBytes; pr.recAt = recs;
        pr.mirror = mirrors ? mirrors[i] : MirrorRange{nullptr, nullptr, nullptr, 0u};
        readBytes += read_size(worlds[out].bodies.size());
        recs += rb;
        sr.pairs.push_back(std::move(pr));
    }
    swapRecords = recs;
    if (sr.pairs.empty()) return BP_OK;
    // The published state of every resident world of the round, for its host copy (transactions
    // may have changed its list): its descriptor, its records (every field) and its bodies' state
    // on both sides. One kernel copies the descriptors and body arrays straight into pinned host
    // memory, one gathers the records into the transfer buffer, which goes down in one copy.
    const size_t copies = 6 * sr.pairs.size(), movesAt = align16(copies * sizeof(bpd::CopyOp));
    swapReadAt = movesAt + align16(sr.pairs.size() * sizeof(bpd::WorldMove));
    bp_status s = ensure_pinned_blob(swapReadAt + readBytes);
    if (s != BP_OK) return s;
    char *rbDev = static_cast<char *>(pinnedBlobDev) + swapReadAt;
    bpd::CopyOp *ops = static_cast<bpd::CopyOp *>(pinnedBlob);
    bpd::WorldMove *moves = reinterpret_cast<bpd::WorldMove *>(static_cast<char *>(pinnedBlob) + movesAt);
    uint32_t nops = 0, maxRecords = 0;
    auto copyOut = [&](const void *src, size_t dstAt, size_t bytes) {
        if (bytes) ops[nops++] = bpd::CopyOp{(uint64_t)(uintptr_t)src, rbDev + dstAt, (uint32_t)bytes, 0u};
    };
    for (size_t j = 0; j < sr.pairs.size(); ++j) {
        const SwapPair &pr = sr.pairs[j];
        const uint32_t out = pr.out, nb = (uint32_t)worlds[out].bodies.size();
        const uint32_t bo = hostDescs[out].bodyOffset;      // a world's ranges do not move while it is resident
        const size_t at = pr.readAt, poseAt = at + align16(sizeof(bpd::WorldDesc));
        copyOut(descs + out, at, sizeof(bpd::WorldDesc));
        copyOut(arena.bodyPose[0] + bo, poseAt, nb * sizeof(float4));
        copyOut(arena.bodyPose[1] + bo, poseAt + nb * sizeof(float4), nb * sizeof(float4));
        copyOut(arena.bodyMotion + bo, poseAt + 2 * nb * sizeof(float4), nb * sizeof(float4));
        copyOut(arena.bodyPrev[0] + bo, poseAt + 3 * nb * sizeof(float4), nb * sizeof(float2));
        copyOut(arena.bodyPrev[1] + bo, poseAt + 3 * nb * sizeof(float4) + nb * sizeof(float2), nb * sizeof(float2));
        moves[j] = bpd::WorldMove{descs + out, xferBubbles + pr.recAt, nullptr, 0u, regions[out].bubbles};
        maxRecords = std::max(maxRecords, regions[out].bubbles);
    }
    BP_HIP(bp_launch_multi_copy(nullptr, static_cast<const bpd::CopyOp *>(pinnedBlobDev), nops, stream));
    BP_HIP(bp_launch_gather_worlds_full(arenaDev, reinterpret_cast<const bpd::WorldMove *>(static_cast<char *>(pinnedBlobDev) + movesAt),
                                        (uint32_t)sr.pairs.size(), maxRecords, stream));
    BP_HIP(hipMemcpyAsync(pinnedRecs, xferBubbles, (size_t)recs * sizeof(bp_bubble), hipMemcpyDeviceToHost, stream));
    return BP_OK;
}

bp_status bp_batch::swap_upload(bp_status *results, hipStream_t stream) {
    SwapRound &sr = swap_round();
    sr.upload.clear();
    swapUploads = 0;
    const char *rb = static_cast<const char *>(pinnedBlob) + swapReadAt;
    for (SwapPair &pr : sr.pairs) {
        HostWorld &o = worlds[pr.out];
        const HostWorld &w = worlds[pr.in];
        const Region &r = regions[pr.out];
        const size_t nb = o.bodies.size();
        const char *at = rb + pr.readAt;
        const bpd::WorldDesc d = *reinterpret_cast<const bpd::WorldDesc *>(at);
        if (d.failed == bpd::kWorldHeld) { results[pr.result] = BP_ERR_UNAVAILABLE; continue; }
        const float4 *pose0 = reinterpret_cast<const float4 *>(at + align16(sizeof(bpd::WorldDesc)));
        const float4 *pose1 = pose0 + nb, *motion = pose1 + nb;
        const float2 *prev0 = reinterpret_cast<const float2 *>(motion + nb), *prev1 = prev0 + nb;
        const uint32_t n = std::min(d.bubbleCount, r.bubbles), p = d.pub;
        o.bubbles.assign(pinnedRecs + pr.recAt, pinnedRecs + pr.recAt + n);
        for (size_t k = 0; k < nb; ++k) {
            bp_body &b = o.bodies[k];
            const float4 pose = p ? pose1[k] : pose0[k];
            const float2 prev = p ? prev1[k] : prev0[k];
            b.pose[0] = pose.x; b.pose[1] = pose.y; b.pose[2] = pose.z;
            b.previous[0] = prev.x; b.previous[1] = prev.y; b.previous[2] = pose.w;
            b.velocity[0] = motion[k].x; b.velocity[1] = motion[k].y; b.velocity[2] = motion[k].z;
        }
        o.failed = d.failed == bpd::kWorldFailed;
        // `in` as pack() would write a world into these ranges (published side 0).
        const uint32_t m = (uint32_t)w.bubbles.size();
        bpd::WorldDesc &dn = pr.dn;
        dn = d;
        dn.bubbleCount = m;
        dn.segmentCount = (uint32_t)w.segments.size(); dn.bodyCount = (uint32_t)w.bodies.size();
        dn.vertexCount = (uint32_t)(w.vertices.size() / 2); dn.triangleCount = (uint32_t)(w.triangles.size() / 4);
        dn.motorCount = (uint32_t)w.motors.size(); dn.bandCount = (uint32_t)w.bands.size(); dn.fieldCount = (uint32_t)w.fields.size();
        dn.coopCellOffset = dn.coopBuckets = dn.coopScratchOffset = 0;
        dn.historyOffset = pr.in * bpd::kFrameLimit;
        dn.pub = 0;
        dn.bubbleCapacity = r.bubbles;
        dn.adjCapacity = r.adjacency;
        dn.mirror = pr.mirror.mirror; dn.bodyMirror = pr.mirror.bodyMirror; dn.mirrorIdx = pr.mirror.mirrorIdx;
        dn.mirrorCapacity = pr.mirror.capacity;
        desc_controls(w, m, dn);
        stage_geometry(w, pr.gs);
        pr.dOut = d;                                         // the device's record: its failure flag is current
        pr.dOut.bubbleCount = 0;
        pr.dOut.mirror = nullptr; pr.dOut.bodyMirror = nullptr; pr.dOut.mirrorIdx = nullptr; pr.dOut.mirrorCapacity = 0;
        sr.upload.push_back(std::move(pr));
    }
    swapUploads = (uint32_t)sr.upload.size();
    if (sr.upload.empty()) return BP_OK;
    // Every world moved in goes up in one blob (records, indices, geometry, both descriptors of
    // each pair and the list of copies); one kernel copies the pieces to their places and one
    // writes each world's records into its structure-of-arrays ranges.
    size_t blobBytes = 0;
    auto room = [&](size_t size) { const size_t at = blobBytes; blobBytes += align16(size); return at; };
    struct Piece { void *dst; const void *src; size_t size; size_t at; };
    std::vector<Piece> pieces;
    pieces.reserve(sr.upload.size() * 24);
    auto piece = [&](void *dst, const void *src, size_t size) {
        if (size) pieces.push_back(Piece{dst, src, size, room(size)});
    };
    auto vec = [&](auto *dst, const auto &src) { piece(dst, src.data(), src.size() * sizeof(src[0])); };
    for (SwapPair &pr : sr.upload) {
        const uint32_t m = (uint32_t)worlds[pr.in].bubbles.size();
        const bpd::WorldDesc &dn = pr.dn;
        const GeometryStage &gs = pr.gs;
        pr.recsAt = room((size_t)m * sizeof(bp_bubble));
        pr.idxAt = room(2 * (size_t)m * sizeof(int32_t));
        vec(arena.segAB + dn.segmentOffset, gs.segAB); vec(arena.segRegion + dn.segmentOffset, gs.segRegion);
        vec(arena.segFlags + dn.segmentOffset, gs.segFlags); vec(arena.segCat + dn.segmentOffset, gs.segCat);
        vec(arena.segMask + dn.segmentOffset, gs.segMask); vec(arena.segId + dn.segmentOffset, gs.segId);
        vec(arena.segOwner + dn.segmentOffset, gs.segOwner);
        vec(arena.bodyPose[0] + dn.bodyOffset, gs.bodyPose); vec(arena.bodyPrev[0] + dn.bodyOffset, gs.bodyPrev);
        vec(arena.bodyMotion + dn.bodyOffset, gs.bodyMotion); vec(arena.bodyShape + dn.bodyOffset, gs.bodyShape);
        vec(arena.bodyLimits + dn.bodyOffset, gs.bodyLimits); vec(arena.bodyRanges + dn.bodyOffset, gs.bodyRanges);
        vec(arena.bodyKind + dn.bodyOffset, gs.bodyKind); vec(arena.bodyId + dn.bodyOffset, gs.bodyId);
        vec(arena.verts + dn.vertexOffset, gs.verts); vec(arena.tris + dn.triangleOffset, gs.tris);
        vec(arena.motors + dn.motorOffset, gs.motors); vec(arena.bands + dn.bandOffset, gs.bands);
        vec(arena.fields + dn.fieldOffset, gs.fields); vec(arena.fieldVerts + dn.fieldVertexOffset, gs.fieldVerts);
        vec(arena.fieldIds + dn.fieldIdOffset, gs.fieldIds);
        piece(descs + pr.in, &pr.dn, sizeof pr.dn);
        piece(descs + pr.out, &pr.dOut, sizeof pr.dOut);
    }
    const size_t opsAt = room(pieces.size() * sizeof(bpd::CopyOp));
    const size_t movesAt = room(sr.upload.size() * sizeof(bpd::WorldMove));
    bp_status s = ensure_pinned_blob(blobBytes);
    if (s != BP_OK) return s;
    if (deviceBlobBytes < blobBytes) {
        if ((s = alloc(deviceBlob, blobBytes)) != BP_OK) return s;
        deviceBlobBytes = blobBytes;
    }
    char *hb = static_cast<char *>(pinnedBlob);
    for (const SwapPair &pr : sr.upload) {
        const HostWorld &w = worlds[pr.in];
        const uint32_t m = (uint32_t)w.bubbles.size();
        if (!m) continue;
        std::memcpy(hb + pr.recsAt, w.bubbles.data(), m * sizeof(bp_bubble));
        motor_band_indices(pr.in, w.bubbles.data(), m, reinterpret_cast<int32_t *>(hb + pr.idxAt));
    }
    bpd::CopyOp *op = reinterpret_cast<bpd::CopyOp *>(hb + opsAt);
    for (const Piece &pc : pieces) {
        std::memcpy(hb + pc.at, pc.src, pc.size);
        *op++ = bpd::CopyOp{pc.at, pc.dst, (uint32_t)pc.size, 0u};
    }
    bpd::WorldMove *mv = reinterpret_cast<bpd::WorldMove *>(hb + movesAt);
    uint32_t nmoves = 0, maxMoved = 0;
    for (const SwapPair &pr : sr.upload) {
        const uint32_t m = (uint32_t)worlds[pr.in].bubbles.size();
        if (!m) continue;
        mv[nmoves++] = bpd::WorldMove{nullptr, deviceBlob + pr.recsAt, reinterpret_cast<const int32_t *>(deviceBlob + pr.idxAt),
                                      pr.dn.bubbleOffset, m};
        maxMoved = std::max(maxMoved, m);
    }
    BP_HIP(hipMemcpyAsync(deviceBlob, hb, blobBytes, hipMemcpyHostToDevice, stream));
    BP_HIP(bp_launch_multi_copy(deviceBlob, reinterpret_cast<const bpd::CopyOp *>(deviceBlob + opsAt), (uint32_t)pieces.size(), stream));
    BP_HIP(bp_launch_scatter_worlds(arenaDev, reinterpret_cast<const bpd::WorldMove *>(deviceBlob + movesAt), nmoves, maxMoved, stream));
    return BP_OK;
}

void bp_batch::swap_finish() {
    SwapRound &sr = swap_round();
    // Residency, ranges, kernel classes (the batch's own class lists), graphs.
    for (const SwapPair &pr : sr.upload) {
        const uint32_t out = pr.out, in = pr.in;
        HostWorld &o = worlds[out], &w = worlds[in];
        const Region r = regions[out];
        hostDescs[in] = pr.dn;
        hostDescs[out] = pr.dOut;
        regions[in] = r;
        regions[out] = Region{};
        const int oldClass = o.packedClass, newClass = default_class_for(w, false);
        w.resident = true; w.capacity = r.bubbles; w.packedClass = newClass;
        o.resident = false; o.capacity = 0; o.packedClass = -1;
        auto &from = classWorlds[oldClass];
        from.erase(std::remove(from.begin(), from.end(), out), from.end());
        classWorlds[newClass].push_back(in);
        if (workEstimate.size() == worlds.size()) workEstimate[in] = (uint32_t)w.bubbles.size();
        if (coopGroups.size() == worlds.size()) { coopGroups[in] = 0; coopGroups[out] = 0; }
    }
    if (!sr.upload.empty()) {
        orderDirty = true;                           // the class lists go up before the next batch launch
        for (auto &g : poolGroups) g = 0;
        release_graph();
    }
    sr.pairs.clear();
    sr.upload.clear();
}


void bp_batch::free_swap_round() {
    delete swapRound;
    swapRound = nullptr;
}

// Queues `steps` steps of the listed worlds (every resident world when `subset` is null) on the
// batch's streams and returns without waiting; finish() collects them.
bp_status bp_batch::launch(const uint32_t *subset, uint32_t subsetCount, uint32_t steps) {
    if (!packed) {
        bp_status s = pack();
        if (s != BP_OK) return s;
    }
    const size_t W = worlds.size();
    pendingSteps = steps;
    pendingList.clear();
    if (subset) {
        std::vector<uint8_t> member(W, 0);
        for (uint32_t j = 0; j < subsetCount; ++j) {
            const uint32_t wi = subset[j];
            if (wi >= W || member[wi]) { set_error("world list holds an unknown or repeated world"); return BP_ERR_INVALID_ARGUMENT; }
            if (!worlds[wi].resident) { set_error("world is not resident"); return BP_ERR_UNAVAILABLE; }
            member[wi] = 1;
            pendingList.push_back(wi);
        }
        // Per class, the listed worlds in the class's longest-first order.
        subsetHost.clear();
        for (int k = 0; k < kClassCount; ++k) {
            launchOffset[k] = (uint32_t)subsetHost.size();
            for (uint32_t wi : classWorlds[k])
                if (member[wi]) subsetHost.push_back(wi);
            launchCount[k] = (uint32_t)subsetHost.size() - launchOffset[k];
        }
        if (!subsetHost.empty())
            BP_HIP(hipMemcpy(subsetOrders, subsetHost.data(), subsetHost.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
        for (size_t wi = 0; wi < W; ++wi) pendingList.push_back((uint32_t)wi);
        for (int k = 0; k < kClassCount; ++k) { launchOffset[k] = 0; launchCount[k] = (uint32_t)classWorlds[k].size(); }
        // Longest-processing-time order by measured device time of the last step: a small world
        // with hundreds of substeps per step starts first instead of last, and its class is
        // launched first.
        if (orderDirty) {
            bp_status s = refresh_order();
            if (s != BP_OK) return s;
        }
    }
    if (!settling)
        for (uint32_t wi : pendingList) worlds[wi].removedIds.clear();
    auto orderOf = [&](int k) -> const uint32_t * { return subset ? subsetOrders + launchOffset[k] : orders[k]; };
    auto worldOf = [&](int k, uint32_t j) -> uint32_t { return subset ? subsetHost[launchOffset[k] + j] : classWorlds[k][j]; };
    if (W == 0 || steps == 0) return BP_OK;
    size_t need = (size_t)steps * W;
    if (statsCapacity < need) {
        bp_status s = alloc(stats, need);
        if (s != BP_OK) return s;
        statsCapacity = need;
    }
    BP_HIP(hipMemsetAsync(stats, 0, need * sizeof(bpd::WorldStatus), streams[0]));
    const bool pooled = !subset && std::any_of(poolGroups, poolGroups + kClassCount, [](uint32_t g) { return g != 0; });
    if (pooled) {
        hipError_t e = bp_launch_fill_u32(cursors, kClassCount + 1, 0u, streams[0]);
        if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    }
    BP_HIP(hipEventRecord(startEvent, streams[0]));
    for (int k = 1; k < kStreams; ++k) BP_HIP(hipStreamWaitEvent(streams[k], startEvent, 0));
    {
        bp_status s = launch_classes(pooled, steps, subset != nullptr, true);
        if (s != BP_OK) return s;
    }
    // Cooperative worlds last, alone on the device: a cooperative grid needs all its workgroups
    // resident at once, so it never shares the device with the other classes' launches.
    if (launchCount[kCoopClass]) {
        hipStream_t stream = streams[0];
        for (int k = 1; k < kStreams; ++k) {
            BP_HIP(hipEventRecord(endEvents[k], streams[k]));
            BP_HIP(hipStreamWaitEvent(stream, endEvents[k], 0));
        }
        BP_HIP(hipEventRecord(classStart[kCoopClass], stream));
        for (uint32_t st = 0; st < steps; ++st)
            for (uint32_t j = 0; j < launchCount[kCoopClass]; ++j) {
                const uint32_t wi = worldOf(kCoopClass, j);
                hipError_t e = bp_launch_world_coop(arenaDev, descs, stats + (size_t)st * W, orderOf(kCoopClass) + j, wi,
                                                    coopGroups[wi], stream);
                if (e != hipSuccess) { set_error(std::string("cooperative launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
            }
        BP_HIP(hipEventRecord(classEnd[kCoopClass], stream));
    }
    for (int k = 0; k < kStreams; ++k) BP_HIP(hipEventRecord(endEvents[k], streams[k]));
    pending = true;
    return BP_OK;
}

// The longest-first order inside every class (uploaded) and across classes, from the measured
// device times, and the pools planned from them.
bp_status bp_batch::refresh_order() {
    for (int k = 0; k < kClassCount; ++k) {
        // Every non-empty list is uploaded: a class's membership may have changed (the autotuner
        // moves worlds between classes and marks the order dirty).
        auto &list = classWorlds[k];
        if (list.empty()) continue;
        std::stable_sort(list.begin(), list.end(), [&](uint32_t a, uint32_t b) { return workEstimate[a] > workEstimate[b]; });
        bp_status s = upload(orders[k], list);
        if (s != BP_OK) return s;
    }
    uint64_t longest[kClassCount];
    for (int k = 0; k < kClassCount; ++k) {
        classOrder[k] = k;
        longest[k] = classWorlds[k].empty() ? 0 : workEstimate[classWorlds[k].front()];
    }
    std::stable_sort(classOrder, classOrder + kClassCount, [&](int a, int b) { return longest[a] > longest[b]; });
    orderDirty = false;
    plan_pools();
    return BP_OK;
}

// The class launches of `steps` steps (every class but the cooperative one) on the classes'
// streams, which already wait for the start. With pools, the pooled classes go first and the
// others wait behind a gate until every pooled workgroup has started (or 2 ms have passed): the
// pools' workgroups need whole CUs, and the small workgroups of the other classes would otherwise
// be scattered over the CUs before the pools are placed. `timed` records the class events.
bp_status bp_batch::launch_classes(bool pooled, uint32_t steps, bool subset, bool timed) {
    const size_t W = worlds.size();
    auto orderOf = [&](int k) -> const uint32_t * { return subset ? subsetOrders + launchOffset[k] : orders[k]; };
    int sequence[kClassCount], pooledClasses = 0, listed = 0;
    uint32_t poolTarget = 0;
    for (int idx = 0; idx < kClassCount; ++idx)
        if (pooled && poolGroups[classOrder[idx]] && launchCount[classOrder[idx]]) {
            sequence[listed++] = classOrder[idx];
            poolTarget += poolGroups[classOrder[idx]];
        }
    pooledClasses = listed;
    for (int idx = 0; idx < kClassCount; ++idx)
        if (!(pooled && poolGroups[classOrder[idx]] && launchCount[classOrder[idx]])) sequence[listed++] = classOrder[idx];
    bool gated = false;
    for (int idx = 0; idx < kClassCount; ++idx) {
        const int k = sequence[idx];
        const uint32_t count = launchCount[k];
        if (!count || k == kCoopClass) continue;
        hipStream_t stream = streams[k % kStreams];
        if (pooledClasses && idx >= pooledClasses) {
            if (!gated) {
                hipError_t e = bp_launch_gate(cursors + kClassCount, poolTarget, 200000u, stream);
                if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
                BP_HIP(hipEventRecord(gateEvent, stream));
                gated = true;
            } else {
                BP_HIP(hipStreamWaitEvent(stream, gateEvent, 0));
            }
        }
        if (timed) BP_HIP(hipEventRecord(classStart[k], stream));
        // An LDS configuration runs all the steps in one launch: each workgroup carries its world
        // through them without waiting for the slowest world of the class at every step.
        if (k == kGlobalClass) {
            for (uint32_t st = 0; st < steps; ++st) {
                hipError_t e = bp_launch_world_step(1024, arenaDev, descs, stats + (size_t)st * W, orderOf(k), count, 0u, stream);
                if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
            }
        } else {
            hipError_t e = bp_launch_world_lds(k, arenaDev, descs, stats, orderOf(k), count, steps, (uint32_t)W, 0u,
                                               pooled ? poolGroups[k] : 0u, cursors + k, cursors + kClassCount, stream);
            if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
        }
        if (timed) BP_HIP(hipEventRecord(classEnd[k], stream));
    }
    return BP_OK;
}

// Collects the launches of launch(): infos[step * listed + j] for the j-th listed world. Without
// `wait`, returns BP_ERR_FRAME_IN_FLIGHT while the device is still working.
bp_status bp_batch::finish(bp_frame_info *infos, bool wait) {
    if (!pending) return BP_OK;
    for (int k = 0; k < kStreams; ++k) {
        if (wait) {
            BP_HIP(hipEventSynchronize(endEvents[k]));
        } else {
            hipError_t e = hipEventQuery(endEvents[k]);
            if (e == hipErrorNotReady) return BP_ERR_FRAME_IN_FLIGHT;
            if (e != hipSuccess) { set_error(std::string("hipEventQuery: ") + hipGetErrorString(e)); pending = false; return BP_ERR_COMMAND_FAILED; }
        }
    }
    pending = false;
    const size_t W = worlds.size();
    const uint32_t steps = pendingSteps;
    if (W == 0 || steps == 0) return BP_OK;
    float ms = 0.0f;
    for (int k = 0; k < kStreams; ++k) {
        float e = 0.0f;
        BP_HIP(hipEventElapsedTime(&e, startEvent, endEvents[k]));
        ms = std::max(ms, e);
    }
    for (int k = 0; k < kClassCount; ++k) {
        lastClassMs[k] = 0.0f;
        if (!launchCount[k]) continue;
        BP_HIP(hipEventElapsedTime(&lastClassMs[k], classStart[k], classEnd[k]));
    }
    bp_status s = collect(steps, infos, pendingList, ms);
    if (s != BP_OK || settling) return s;
    const std::vector<uint32_t> listed = pendingList;
    return settle_held(steps, infos, listed);
}

// Deletes the region-removed bubbles of the held worlds (one workgroup each) and records their ids.
// removals[j]: bubbles the last step of held[j] removed.
bp_status bp_batch::drop_removed(const std::vector<uint32_t> &held, const std::vector<uint32_t> &removals) {
    const uint32_t count = (uint32_t)held.size();
    if (!count) return BP_OK;
    std::vector<bpd::DropRecord> recs(count);
    uint32_t total = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t wi = held[j];
        const HostWorld &w = worlds[wi];
        const uint32_t n = hostDescs[wi].bubbleCount >= removals[j] ? hostDescs[wi].bubbleCount - removals[j] : 0u;
        recs[j].world = wi;
        recs[j].buckets = w.config.grid_bucket_count == 0 ? auto_buckets(n) : std::min(w.config.grid_bucket_count, bpd::kMaxBuckets);
        recs[j].removedOffset = total;
        recs[j].reserved = 0;
        total += removals[j];
    }
    if (dropCapacity < count) { bp_status s = alloc(dropRecs, count); if (s != BP_OK) return s; dropCapacity = count; }
    if (dropIdCapacity < std::max<uint32_t>(total, 1)) {
        bp_status s = alloc(dropIds, std::max<uint32_t>(total, 1));
        if (s != BP_OK) return s;
        dropIdCapacity = std::max<uint32_t>(total, 1);
    }
    BP_HIP(hipMemcpyAsync(dropRecs, recs.data(), count * sizeof(bpd::DropRecord), hipMemcpyHostToDevice, streams[0]));
    hipError_t e = bp_launch_drop_removed(arenaDev, descs, dropRecs, count, dropIds, streams[0]);
    if (e != hipSuccess) { set_error(std::string("drop kernel: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    std::vector<int64_t> ids(total);
    if (total) BP_HIP(hipMemcpyAsync(ids.data(), dropIds, total * sizeof(int64_t), hipMemcpyDeviceToHost, streams[0]));
    BP_HIP(hipStreamSynchronize(streams[0]));
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t wi = held[j];
        HostWorld &w = worlds[wi];
        auto &bs = w.bubbles;
        for (uint32_t k = 0; k < removals[j]; ++k) {
            const int64_t id = ids[recs[j].removedOffset + k];
            w.removedIds.push_back(id);
            auto it = std::lower_bound(bs.begin(), bs.end(), id, [](const bp_bubble &b, int64_t v) { return b.id < v; });
            if (it != bs.end() && it->id == id) bs.erase(it);
        }
        hostDescs[wi].bubbleCount = (uint32_t)bs.size();
        hostDescs[wi].buckets = recs[j].buckets;
        hostDescs[wi].failed = bpd::kWorldRunnable;
    }
    return BP_OK;
}

// After a launch: worlds whose last executed step removed bubbles from the region are held on the
// device. A batch deletes their removed bubbles and runs the launch's remaining steps for them,
// writing their results into infos (step-major over `listed`, may be null); a plain world only
// makes them runnable again.
bp_status bp_batch::settle_held(uint32_t steps, bp_frame_info *infos, const std::vector<uint32_t> &listed) {
    const size_t W = worlds.size();
    std::vector<uint32_t> done(listed.size(), 0), held, removals;
    std::vector<size_t> slot;                         // position of a held world in `listed`
    for (size_t j = 0; j < listed.size(); ++j) {
        const uint32_t wi = listed[j];
        for (uint32_t st = 0; st < steps; ++st) {
            const bpd::WorldStatus &r = hostStats[(size_t)st * W + wi];
            if (!r.executed) break;
            done[j] = st + 1;
        }
        if (!done[j]) continue;
        const bpd::WorldStatus &last = hostStats[(size_t)(done[j] - 1) * W + wi];
        if (last.status == BP_OK && last.removals) { held.push_back(wi); removals.push_back(last.removals); slot.push_back(j); }
    }
    if (held.empty()) return BP_OK;
    if (!dropOnHold) {
        for (uint32_t wi : held) {
            bpd::WorldDesc d;
            BP_HIP(hipMemcpy(&d, descs + wi, sizeof d, hipMemcpyDeviceToHost));
            if (d.failed == bpd::kWorldHeld) { d.failed = bpd::kWorldRunnable; BP_HIP(hipMemcpy(descs + wi, &d, sizeof d, hipMemcpyHostToDevice)); }
        }
        return BP_OK;
    }
    settling = true;
    bp_status s = BP_OK;
    const size_t L = listed.size();
    while (!held.empty() && s == BP_OK) {
        s = drop_removed(held, removals);
        if (s != BP_OK) break;
        // Worlds with steps left, grouped by how many.
        std::vector<uint32_t> nextHeld, nextRemovals;
        std::vector<size_t> nextSlot;
        std::vector<size_t> order(held.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = k;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return done[slot[a]] < done[slot[b]]; });
        for (size_t a = 0; a < order.size() && s == BP_OK;) {
            const uint32_t already = done[slot[order[a]]];
            size_t b = a;
            std::vector<uint32_t> group;
            std::vector<size_t> groupSlot;
            while (b < order.size() && done[slot[order[b]]] == already) {
                group.push_back(held[order[b]]);
                groupSlot.push_back(slot[order[b]]);
                ++b;
            }
            a = b;
            if (already >= steps) continue;
            const uint32_t rest = steps - already;
            std::vector<bp_frame_info> tmp((size_t)rest * group.size());
            s = launch(group.data(), (uint32_t)group.size(), rest);
            if (s == BP_OK) s = finish(tmp.data(), true);
            if (s != BP_OK) break;
            for (size_t g = 0; g < group.size(); ++g) {
                const size_t j = groupSlot[g];
                uint32_t ran = 0;
                for (uint32_t st = 0; st < rest; ++st) {
                    const bp_frame_info &r = tmp[(size_t)st * group.size() + g];
                    if (infos) infos[(size_t)(already + st) * L + j] = r;
                    if (hostStats[(size_t)st * W + group[g]].executed) ran = st + 1;
                }
                done[j] = already + ran;
                if (ran && tmp[(size_t)(ran - 1) * group.size() + g].status == BP_OK &&
                    tmp[(size_t)(ran - 1) * group.size() + g].region_removals) {
                    nextHeld.push_back(group[g]);
                    nextRemovals.push_back(tmp[(size_t)(ran - 1) * group.size() + g].region_removals);
                    nextSlot.push_back(j);
                }
            }
        }
        held.swap(nextHeld); removals.swap(nextRemovals); slot.swap(nextSlot);
    }
    settling = false;
    return s;
}

// Result records of the last `steps` steps into infos[step * listed + j] for the j-th listed world.
bp_status bp_batch::collect(uint32_t steps, bp_frame_info *infos, const std::vector<uint32_t> &listedWorlds, float ms) {
    const size_t W = worlds.size();
    lastGpuMs = ms;
    const size_t need = (size_t)steps * W;
    hostStats.resize(need);
    BP_HIP(hipMemcpy(hostStats.data(), stats, need * sizeof(bpd::WorldStatus), hipMemcpyDeviceToHost));
    if (workEstimate.size() == W)
        for (uint32_t wi : listedWorlds) {
            // Device ticks per executed step; the order changes only when an estimate moves by
            // more than a quarter (the ticks of a world vary a little from step to step).
            uint64_t ticks = 0, executed = 0;
            for (uint32_t st = 0; st < steps; ++st) {
                const bpd::WorldStatus &r = hostStats[(size_t)st * W + wi];
                if (r.executed) { ticks += r.ticks; ++executed; }
            }
            if (!executed) continue;
            const uint64_t est = std::max<uint64_t>(1, ticks / executed), old = workEstimate[wi];
            if (4 * est > 5 * old || 5 * est < 4 * old) { workEstimate[wi] = est; orderDirty = true; }
        }
    if (workEstimate.size() == W && listedWorlds.size() == W) estimatesMeasured = true;
    if (!infos) return BP_OK;
    const size_t listed = listedWorlds.size();
    for (uint32_t st = 0; st < steps; ++st)
        for (size_t j = 0; j < listed; ++j) {
            const uint32_t wi = listedWorlds[j];
            bp_frame_info &out = infos[(size_t)st * listed + j];
            fill_info(hostStats[(size_t)st * W + wi], &out);
            const int cls = worlds[wi].packedClass;
            out.kernel_class = cls < 0 ? 0xffffffffu : (uint32_t)cls;
            if (cls == kGlobalClass) out.execution_path = BP_PATH_WORKGROUP;
            if (cls == kCoopClass) out.execution_path = BP_PATH_COOPERATIVE;
            out.gpu_milliseconds = ms / steps;
            const bpd::WorldStatus &rec = hostStats[(size_t)st * W + wi];
            if (!rec.executed) out.status = BP_ERR_UNAVAILABLE;
            // A record that claims success must describe a finished step within the substep limit
            // (the read-back consistency check of the reference's frame runner): otherwise the frame
            // is invalid and nothing it reports may be used.
            else if (rec.status == BP_OK && (rec.substeps > bpd::kFrameLimit || rec.time[0] - rec.time[1] > 0.0f ||
                                             rec.removals + rec.wraps > hostDescs[wi].bubbleCapacity))
                out.status = BP_ERR_INVALID_FRAME;
        }
    return BP_OK;
}

// Region results of a world: the ids the region removed (a batch: during its last stepping call,
// after which they were deleted; a plain world: in its last published step, whose frame still
// holds them) and the ids it wrapped in the world's last published step, ascending.
bp_status bp_batch::region_results(uint32_t wi, int64_t *removed, uint32_t removedCapacity, uint32_t *removedCount,
                                   int64_t *wrapped, uint32_t wrappedCapacity, uint32_t *wrappedCount) {
    std::vector<int64_t> rm, wr;
    const HostWorld &w = worlds[wi];
    if (packed && deviceValid && w.resident) {
        bpd::WorldDesc d;
        BP_HIP(hipMemcpy(&d, descs + wi, sizeof d, hipMemcpyDeviceToHost));
        std::vector<uint32_t> region(d.bubbleCount);
        std::vector<int64_t> ids(d.bubbleCount);
        if (d.bubbleCount) {
            BP_HIP(hipMemcpy(region.data(), arena.region[d.pub] + d.bubbleOffset, d.bubbleCount * sizeof(uint32_t), hipMemcpyDeviceToHost));
            BP_HIP(hipMemcpy(ids.data(), arena.id + d.bubbleOffset, d.bubbleCount * sizeof(int64_t), hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < d.bubbleCount; ++i) {
            if (region[i] == BP_REGION_REMOVED) rm.push_back(ids[i]);
            if (region[i] == BP_REGION_WRAPPED) wr.push_back(ids[i]);
        }
    }
    if (dropOnHold) rm = w.removedIds;
    std::sort(rm.begin(), rm.end());
    if (removedCount) *removedCount = (uint32_t)rm.size();
    if (wrappedCount) *wrappedCount = (uint32_t)wr.size();
    if ((rm.size() && removed && removedCapacity < rm.size()) || (wr.size() && wrapped && wrappedCapacity < wr.size()))
        return BP_ERR_CAPACITY_OVERFLOW;
    if (removed) std::copy(rm.begin(), rm.end(), removed);
    if (wrapped) std::copy(wr.begin(), wr.end(), wrapped);
    return BP_OK;
}

// Deletes bubbles by id from a world's published state (the next snapshot of a public world that
// dropped them, NativePublicWorld.swift:264-267), with their motor and band records.
bp_status bp_batch::remove_bubbles(uint32_t wi, const int64_t *ids, uint32_t count) {
    bp_status s = download_state();
    if (s != BP_OK) return s;
    HostWorld &w = worlds[wi];
    std::vector<int64_t> gone(ids, ids + count);
    std::sort(gone.begin(), gone.end());
    for (int64_t id : gone) {
        auto it = std::lower_bound(w.bubbles.begin(), w.bubbles.end(), id, [](const bp_bubble &b, int64_t v) { return b.id < v; });
        if (it == w.bubbles.end() || it->id != id) { set_error("removal names a bubble the world does not hold"); return BP_ERR_INVALID_SNAPSHOT; }
    }
    auto isGone = [&](int64_t id) { return std::binary_search(gone.begin(), gone.end(), id); };
    w.bubbles.erase(std::remove_if(w.bubbles.begin(), w.bubbles.end(), [&](const bp_bubble &b) { return isGone(b.id); }), w.bubbles.end());
    w.motors.erase(std::remove_if(w.motors.begin(), w.motors.end(), [&](const bp_motor &m) { return isGone(m.bubble_id); }), w.motors.end());
    w.bands.erase(std::remove_if(w.bands.begin(), w.bands.end(), [&](const bp_band &m) { return isGone(m.bubble_id); }), w.bands.end());
    packed = false;
    return BP_OK;
}

// FP64 audit of every world's published state (audit_kernel, bp_audit.hip).
bp_status bp_batch::audit(bp_audit *out, uint32_t count) {
    if (count != worlds.size() || (count && !out)) return BP_ERR_INVALID_ARGUMENT;
    if (count == 0) return BP_OK;
    if (!packed) {
        bp_status s = pack();
        if (s != BP_OK) return s;
    }
    void *dev = nullptr;
    BP_HIP(hipMalloc(&dev, count * sizeof(bp_audit)));
    hipError_t e = bp_launch_audit(arenaDev, descs, count, dev, streams[0]);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dev, count * sizeof(bp_audit), hipMemcpyDeviceToHost, streams[0]);
    if (e == hipSuccess) e = hipStreamSynchronize(streams[0]);
    (void)hipFree(dev);
    if (e != hipSuccess) { set_error(std::string("audit: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    for (uint32_t wi = 0; wi < count; ++wi)
        if (!worlds[wi].resident) {                    // held in host memory: nothing on the device to audit
            const double nan = std::nan("");
            out[wi] = bp_audit{{nan, nan}, nan, nan, nan, 0u, 0u};
        }
    return BP_OK;
}

// End-of-step contact records of world `wi` (contacts_kernel, bp_contacts.hip): the pair and
// segment records, or with `bodies` the bubble-body records, read back and put body by body.
bp_status bp_batch::contacts(uint32_t wi, bp_contact *out, uint32_t capacity, uint32_t *count, bool bodies) {
    if (wi >= worlds.size() || (capacity && !out) || !count) return BP_ERR_INVALID_ARGUMENT;
    if (!packed) {
        bp_status s = pack();
        if (s != BP_OK) return s;
    }
    if (!worlds[wi].resident) { set_error("world is not resident"); return BP_ERR_UNAVAILABLE; }
    bpd::WorldDesc d;
    BP_HIP(hipMemcpy(&d, descs + wi, sizeof(d), hipMemcpyDeviceToHost));
    if (!d.exportAdjacency) {
        set_error("contact records read the rows of the last grid build: set export_adjacency");
        return BP_ERR_UNAVAILABLE;
    }
    *count = 0;
    if (d.bubbleCount == 0) return BP_OK;
    // Body records are reordered on the host, so all of them are read: a second launch with room
    // for every record when the first finds more than `capacity`.
    uint32_t room = capacity, kept = 0;
    std::vector<bpd::ContactRecord> host;
    for (int pass = 0; pass < 2; ++pass) {
        void *dev = nullptr;
        const size_t offsetBytes = (size_t)(bodies ? 1 : 2) * d.bubbleCount * sizeof(uint32_t);
        const size_t recordBytes = (size_t)room * sizeof(bpd::ContactRecord);
        BP_HIP(hipMalloc(&dev, offsetBytes + recordBytes + 256));
        uint32_t *offsets = static_cast<uint32_t *>(dev);
        auto *records = reinterpret_cast<bpd::ContactRecord *>(static_cast<char *>(dev) + ((offsetBytes + 255) / 256) * 256);
        uint32_t *total = reinterpret_cast<uint32_t *>(static_cast<char *>(dev) + ((offsetBytes + 255) / 256) * 256 + recordBytes);
        hipError_t e = bodies ? bp_launch_body_contacts(arenaDev, descs, wi, offsets, records, room, total, streams[0])
                              : bp_launch_contacts(arenaDev, descs, wi, offsets, records, room, total, streams[0]);
        if (e == hipSuccess) e = hipMemcpyAsync(&kept, total, sizeof(kept), hipMemcpyDeviceToHost, streams[0]);
        if (e == hipSuccess) e = hipStreamSynchronize(streams[0]);
        if (e == hipSuccess && kept) {
            host.resize(std::min(kept, room));
            e = hipMemcpy(host.data(), records, host.size() * sizeof(bpd::ContactRecord), hipMemcpyDeviceToHost);
        }
        (void)hipFree(dev);
        if (e != hipSuccess) { set_error(std::string("contacts: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
        if (!bodies || kept <= room) break;
        room = kept;
    }
    if (bodies)
        std::stable_sort(host.begin(), host.end(),
                         [](const bpd::ContactRecord &a, const bpd::ContactRecord &b) { return a.indexB < b.indexB; });
    if (host.size() > capacity) host.resize(capacity);
    for (size_t k = 0; k < host.size(); ++k) {
        const bpd::ContactRecord &r = host[k];
        bp_contact &c = out[k];
        // The contact id of NativeMetalSolver.swift:1047-1052 (FNV-1a steps over the key).
        uint64_t h = 14695981039346656037ull;
        h = (h ^ (uint64_t)r.kind) * 1099511628211ull;
        h = (h ^ (uint64_t)r.first) * 1099511628211ull;
        h = (h ^ (uint64_t)r.second) * 1099511628211ull;
        c.id = h;
        c.first = r.first; c.second = r.second;
        c.kind = r.kind; c.feature = r.feature;
        c.normal[0] = r.normal[0]; c.normal[1] = r.normal[1];
        c.point[0] = r.point[0]; c.point[1] = r.point[1];
        c.penetration = r.penetration;
        c.allowed_side = r.allowedSide;
        c.impulse = r.impulse;
        c.index_b = r.indexB;
    }
    *count = kept;
    return BP_OK;
}

// Contour stride of a world's bubbles (NativeMetalSolver.swift:407-411): max over bubbles of
// max(64, ceil(4 pi r / maxContourSegmentLength)), in double as the reference computes it.
bp_status bp_batch::contour_stride(uint32_t wi, const bp_contour_config *cfg, uint32_t *stride) {
    if (wi >= worlds.size() || !cfg || !stride) return BP_ERR_INVALID_ARGUMENT;
    if (!(cfg->max_segment_length > 0.0f) || !std::isfinite(cfg->max_segment_length)) {
        set_error("max_segment_length must be finite and positive");
        return BP_ERR_INVALID_ARGUMENT;
    }
    double best = 64.0;
    for (const bp_bubble &b : worlds[wi].bubbles)
        best = std::max(best, std::ceil(4.0 * M_PI * (double)b.radius / (double)cfg->max_segment_length));
    if (!(best <= 4294967295.0)) { set_error("contour stride overflows"); return BP_ERR_ALLOCATION_FAILED; }
    *stride = (uint32_t)best;
    return BP_OK;
}

bp_status bp_batch::contours(uint32_t wi, const bp_contour_config *cfg, uint32_t stride, float *points,
                             bp_bubble_contour *shapes, uint32_t count) {
    if (wi >= worlds.size() || !cfg || !points || !shapes) return BP_ERR_INVALID_ARGUMENT;
    if (!(cfg->max_segment_length > 0.0f) || !std::isfinite(cfg->max_segment_length) || !(cfg->surface_tension >= 0.0f) ||
        !std::isfinite(cfg->surface_tension) || !(cfg->position_tolerance >= 0.0f) || !std::isfinite(cfg->position_tolerance)) {
        set_error("contour configuration: positive segment length, non-negative tension and tolerance");
        return BP_ERR_INVALID_ARGUMENT;
    }
    if (!packed) {
        bp_status s = pack();
        if (s != BP_OK) return s;
    }
    if (!worlds[wi].resident) { set_error("world is not resident"); return BP_ERR_UNAVAILABLE; }
    bpd::WorldDesc d;
    BP_HIP(hipMemcpy(&d, descs + wi, sizeof(d), hipMemcpyDeviceToHost));
    if (count != d.bubbleCount) { set_error("count must equal the world's bubble count"); return BP_ERR_INVALID_ARGUMENT; }
    if (!d.exportAdjacency) {
        set_error("contours read the rows of the last grid build: set export_adjacency");
        return BP_ERR_UNAVAILABLE;
    }
    uint32_t need = 0;
    bp_status s = contour_stride(wi, cfg, &need);
    if (s != BP_OK) return s;
    if (stride < need) { set_error("stride is below the world's contour stride"); return BP_ERR_INVALID_ARGUMENT; }
    if (stride > bp_contour_capacity()) { set_error("contour stride above the kernel's 2,048 samples"); return BP_ERR_CAPACITY_OVERFLOW; }
    if (count == 0) return BP_OK;
    void *dev = nullptr;
    const size_t pointBytes = (size_t)count * stride * sizeof(float2);
    const size_t shapeOffset = (pointBytes + 255) / 256 * 256;
    BP_HIP(hipMalloc(&dev, shapeOffset + (size_t)count * sizeof(bp_bubble_contour)));
    float2 *devPoints = static_cast<float2 *>(dev);
    auto *devShapes = reinterpret_cast<bp_bubble_contour *>(static_cast<char *>(dev) + shapeOffset);
    hipError_t e = bp_launch_contours(arenaDev, descs, wi, count, *cfg, stride, devPoints, devShapes, streams[0]);
    if (e == hipSuccess) e = hipMemcpyAsync(points, devPoints, pointBytes, hipMemcpyDeviceToHost, streams[0]);
    if (e == hipSuccess) e = hipMemcpyAsync(shapes, devShapes, count * sizeof(bp_bubble_contour), hipMemcpyDeviceToHost, streams[0]);
    if (e == hipSuccess) e = hipStreamSynchronize(streams[0]);
    (void)hipFree(dev);
    if (e != hipSuccess) { set_error(std::string("contours: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    return BP_OK;
}

bp_status bp_batch::step(uint32_t steps, bp_frame_info *infos) {
    bp_status s = launch(nullptr, 0, steps);
    if (s != BP_OK) return s;
    return finish(infos, true);
}

// Captures `steps` steps of every resident world into hipGraphs, one per kernel class, each
// captured on the class's stream: the class's launch (behind the gate when pools are planned and
// the class is not pooled). replay() clears the result records and the pool words with fill
// kernels (no memset node), then launches every class's graph on its stream and the cooperative
// class's graph last, as launch() orders direct launches.
bp_status bp_batch::capture(uint32_t steps) {
    if (steps == 0) return BP_ERR_INVALID_ARGUMENT;
    if (!packed) {
        bp_status s = pack();
        if (s != BP_OK) return s;
    }
    release_graph();
    if (orderDirty) {
        bp_status s = refresh_order();
        if (s != BP_OK) return s;
    }
    const size_t W = worlds.size();
    const size_t need = (size_t)steps * W;
    if (statsCapacity < need) {                       // allocations happen before the capture
        bp_status s = alloc(stats, need);
        if (s != BP_OK) return s;
        statsCapacity = need;
    }
    if (!forkEvent) BP_HIP(hipEventCreateWithFlags(&forkEvent, hipEventDisableTiming));
    for (int k = 0; k < kClassCount; ++k) { launchOffset[k] = 0; launchCount[k] = (uint32_t)classWorlds[k].size(); }
    const bool pooled = std::any_of(poolGroups, poolGroups + kClassCount, [](uint32_t g) { return g != 0; });
    uint32_t poolTarget = 0;
    for (int k = 0; k < kClassCount; ++k) if (pooled && launchCount[k]) poolTarget += poolGroups[k];
    for (int k = 0; k < kClassCount; ++k) {
        const uint32_t count = launchCount[k];
        if (!count) continue;
        hipStream_t stream = k == kCoopClass ? streams[0] : streams[k % kStreams];
        BP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        if (k == kCoopClass) {
            for (uint32_t st = 0; st < steps && e == hipSuccess; ++st)
                for (uint32_t j = 0; j < count && e == hipSuccess; ++j) {
                    const uint32_t wi = classWorlds[kCoopClass][j];
                    e = bp_launch_world_coop(arenaDev, descs, stats + (size_t)st * W, orders[kCoopClass] + j, wi, coopGroups[wi], stream);
                }
        } else {
            if (pooled && !poolGroups[k]) e = bp_launch_gate(cursors + kClassCount, poolTarget, 200000u, stream);
            if (k == kGlobalClass) {
                for (uint32_t st = 0; st < steps && e == hipSuccess; ++st)
                    e = bp_launch_world_step(1024, arenaDev, descs, stats + (size_t)st * W, orders[k], count, 0u, stream);
            } else if (e == hipSuccess) {
                e = bp_launch_world_lds(k, arenaDev, descs, stats, orders[k], count, steps, (uint32_t)W, 0u,
                                        pooled ? poolGroups[k] : 0u, cursors + k, cursors + kClassCount, stream);
            }
        }
        hipGraph_t g = nullptr;
        hipError_t end = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess || end != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            release_graph();
            set_error(std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : end));
            return BP_ERR_COMMAND_SETUP_FAILED;
        }
        graphs[k] = g;
        hipError_t ie = hipGraphInstantiate(&graphExecs[k], g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            release_graph();
            set_error(std::string("graph instantiate: ") + hipGetErrorString(ie));
            return BP_ERR_COMMAND_SETUP_FAILED;
        }
    }
    graphSteps = steps;
    graphPooled = pooled;
    return BP_OK;
}

// One launch of the captured graphs, then the result records of their steps for every world.
bp_status bp_batch::replay(bp_frame_info *infos) {
    bool any = false;
    for (auto g : graphExecs) any = any || g;
    if (!any || !packed || !graphSteps) { set_error("no captured graph (capture again after the batch changed)"); return BP_ERR_INVALID_ARGUMENT; }
    const size_t W = worlds.size();
    const size_t need = (size_t)graphSteps * W;
    hipStream_t origin = streams[0];
    BP_HIP(hipEventRecord(startEvent, origin));
    hipError_t e = bp_launch_fill_u32(reinterpret_cast<uint32_t *>(stats), need * sizeof(bpd::WorldStatus) / 4, 0u, origin);
    if (e == hipSuccess) e = bp_launch_fill_u32(cursors, kClassCount + 1, 0u, origin);
    if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipEventRecord(forkEvent, origin));
    for (int k = 1; k < kStreams; ++k) BP_HIP(hipStreamWaitEvent(streams[k], forkEvent, 0));
    // Pooled classes first (their workgroups need whole CUs), then the others (gated inside their graphs).
    for (int pass = 0; pass < 2; ++pass)
        for (int idx = 0; idx < kClassCount; ++idx) {
            const int k = classOrder[idx];
            if (!graphExecs[k] || k == kCoopClass) continue;
            if ((pass == 0) != (graphPooled && poolGroups[k] != 0)) continue;
            BP_HIP(hipEventRecord(classStart[k], streams[k % kStreams]));
            BP_HIP(hipGraphLaunch(graphExecs[k], streams[k % kStreams]));
            BP_HIP(hipEventRecord(classEnd[k], streams[k % kStreams]));
        }
    for (int k = 1; k < kStreams; ++k) {
        BP_HIP(hipEventRecord(endEvents[k], streams[k]));
        BP_HIP(hipStreamWaitEvent(origin, endEvents[k], 0));
    }
    if (graphExecs[kCoopClass]) {
        BP_HIP(hipEventRecord(classStart[kCoopClass], origin));
        BP_HIP(hipGraphLaunch(graphExecs[kCoopClass], origin));
        BP_HIP(hipEventRecord(classEnd[kCoopClass], origin));
    }
    BP_HIP(hipEventRecord(endEvents[0], origin));
    BP_HIP(hipEventSynchronize(endEvents[0]));
    float ms = 0.0f;
    BP_HIP(hipEventElapsedTime(&ms, startEvent, endEvents[0]));
    for (int k = 0; k < kClassCount; ++k) {
        lastClassMs[k] = 0.0f;
        if (graphExecs[k]) BP_HIP(hipEventElapsedTime(&lastClassMs[k], classStart[k], classEnd[k]));
    }
    std::vector<uint32_t> all(W);
    for (uint32_t wi = 0; wi < all.size(); ++wi) all[wi] = wi;
    for (auto &w : worlds) w.removedIds.clear();
    bp_status s = collect(graphSteps, infos, all, ms);
    if (s != BP_OK) return s;
    // A world held for region removals leaves the graph's launches: the deletion moves it to a
    // smaller bubble count, which the captured launches still cover (counts are read from the
    // descriptors), so the graph stays valid.
    return settle_held(graphSteps, infos, all);
}

// ---------------------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------------------
extern "C" {

uint32_t bp_abi_version(void) { return BP_ABI_VERSION; }

static_assert(sizeof(bp_contact) == 64, "bp_contact is 64 bytes");
static_assert(sizeof(bpd::ContactRecord) == 56, "device contact record");

void bp_abi_record_sizes(uint32_t *out) {
    if (!out) return;
    out[0] = sizeof(bp_bubble); out[1] = sizeof(bp_segment); out[2] = sizeof(bp_body); out[3] = sizeof(bp_motor);
    out[4] = sizeof(bp_band); out[5] = sizeof(bp_world_config); out[6] = sizeof(bp_frame_info);
    out[7] = sizeof(bp_batch_world_desc); out[8] = sizeof(bp_field);
}

void bp_world_config_default(bp_world_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->time_step = 1.0f / 60.0f;
    c->contact_tolerance = 0.001f;
    c->separation_tolerance = 0.002f;
    c->candidate_limit = 128;
    c->substep_split = 1;
    c->grid_bucket_count = 0;    // automatic: next power of two >= bubble count, up to 8192
}

const char *bp_status_string(bp_status s) {
    switch (s) {
        case BP_OK: return "ok";
        case BP_ERR_INVALID_ARGUMENT: return "invalid argument";
        case BP_ERR_RESOURCE_UNAVAILABLE: return "resource unavailable";
        case BP_ERR_ALLOCATION_FAILED: return "allocation failed";
        case BP_ERR_COMMAND_SETUP_FAILED: return "command setup failed";
        case BP_ERR_COMMAND_FAILED: return "command failed";
        case BP_ERR_INVALID_SNAPSHOT: return "invalid snapshot";
        case BP_ERR_CAPACITY_OVERFLOW: return "capacity overflow";
        case BP_ERR_FRAME_IN_FLIGHT: return "frame in flight";
        case BP_ERR_NON_FINITE: return "non-finite state";
        case BP_ERR_GUARD_FAILURE: return "guard failure";
        case BP_ERR_CONTACT_VIOLATION: return "contact violation";
        case BP_ERR_SUBSTEP_LIMIT_EXCEEDED: return "substep limit exceeded";
        case BP_ERR_UNAVAILABLE: return "unavailable";
        case BP_ERR_INVALID_FRAME: return "invalid frame";
        default: return "unknown status";
    }
}

const char *bp_last_error(void) { return g_lastError.c_str(); }

static bp_status device_available() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        set_error("no HIP device is available");
        return BP_ERR_RESOURCE_UNAVAILABLE;
    }
    return BP_OK;
}

// ---- batch ----
bp_status bp_batch_create(bp_batch **out) {
    if (!out) return BP_ERR_INVALID_ARGUMENT;
    bp_status s = device_available();
    if (s != BP_OK) return s;
    *out = new bp_batch();
    return BP_OK;
}

void bp_batch_destroy(bp_batch *b) { delete b; }

static bp_status world_from_desc(HostWorld &w, const bp_batch_world_desc *d) {
    if (!d) return BP_ERR_INVALID_ARGUMENT;
    if ((d->bubble_count && !d->bubbles) || (d->segment_count && !d->segments) || (d->body_count && !d->bodies) ||
        (d->vertex_count && !d->vertices) || (d->triangle_count && !d->triangles) || (d->motor_count && !d->motors) ||
        (d->band_count && !d->bands) || (d->field_count && !d->fields) || (d->field_vertex_count && !d->field_vertices) ||
        (d->field_filter_count && !d->field_filter_ids)) {
        set_error("null array with a non-zero count");
        return BP_ERR_INVALID_ARGUMENT;
    }
    w.config = d->config;
    w.bubbles.assign(d->bubbles, d->bubbles + d->bubble_count);
    w.segments.assign(d->segments, d->segments + d->segment_count);
    w.bodies.assign(d->bodies, d->bodies + d->body_count);
    w.vertices.assign(d->vertices, d->vertices + 2 * (size_t)d->vertex_count);
    w.triangles.assign(d->triangles, d->triangles + 4 * (size_t)d->triangle_count);
    w.motors.assign(d->motors, d->motors + d->motor_count);
    w.bands.assign(d->bands, d->bands + d->band_count);
    w.fields.assign(d->fields, d->fields + d->field_count);
    w.fieldVertices.assign(d->field_vertices, d->field_vertices + 2 * (size_t)d->field_vertex_count);
    w.fieldIds.assign(d->field_filter_ids, d->field_filter_ids + d->field_filter_count);
    return validate_world(w);
}

bp_status bp_batch_add_world(bp_batch *b, const bp_batch_world_desc *desc, uint32_t *out_index) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_state();
    if (s != BP_OK) return s;
    HostWorld w;
    s = world_from_desc(w, desc);
    if (s != BP_OK) return s;
    b->worlds.push_back(std::move(w));
    b->packed = false;
    if (out_index) *out_index = (uint32_t)(b->worlds.size() - 1);
    return BP_OK;
}

uint32_t bp_batch_world_count(const bp_batch *b) { return b ? (uint32_t)b->worlds.size() : 0u; }

uint32_t bp_batch_bubble_count(const bp_batch *b, uint32_t world) {
    return b && world < b->worlds.size() ? (uint32_t)b->worlds[world].bubbles.size() : 0u;
}

bp_status bp_batch_step(bp_batch *b, uint32_t steps, bp_frame_info *infos) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(b->mutex, std::try_to_lock);
    if (!lock.owns_lock() || b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    b->inFlight = true;
    bp_status s = b->step(steps, infos);
    b->inFlight = false;
    return s;
}

bp_status bp_batch_step_worlds(bp_batch *b, const uint32_t *worlds, uint32_t count, uint32_t steps, bp_frame_info *infos) {
    if (!b || (count && !worlds)) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(b->mutex, std::try_to_lock);
    if (!lock.owns_lock() || b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    b->inFlight = true;
    bp_status s = b->launch(worlds, count, steps);
    if (s == BP_OK) s = b->finish(infos, true);
    b->inFlight = false;
    return s;
}

bp_status bp_batch_submit_worlds(bp_batch *b, const uint32_t *worlds, uint32_t count, uint32_t steps) {
    if (!b || (count && !worlds)) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(b->mutex, std::try_to_lock);
    if (!lock.owns_lock() || b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->launch(worlds, count, steps);
}

bp_status bp_batch_poll(bp_batch *b, bp_frame_info *infos) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(b->mutex, std::try_to_lock);
    if (!lock.owns_lock() || b->inFlight || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    if (!b->pending) { set_error("nothing was submitted"); return BP_ERR_INVALID_ARGUMENT; }
    return b->finish(infos, false);
}

bp_status bp_batch_write_bubbles(bp_batch *b, uint32_t world, const bp_bubble *bubbles, uint32_t count) {
    if (!b || world >= b->worlds.size() || (count && !bubbles)) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->write_bubbles(world, bubbles, count);
}

bp_status bp_batch_region_results(bp_batch *b, uint32_t world, int64_t *removed, uint32_t removed_capacity,
                                  uint32_t *out_removed, int64_t *wrapped, uint32_t wrapped_capacity, uint32_t *out_wrapped) {
    if (!b || world >= b->worlds.size()) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->region_results(world, removed, removed_capacity, out_removed, wrapped, wrapped_capacity, out_wrapped);
}

bp_status bp_batch_set_resident(bp_batch *b, uint32_t world, uint32_t resident) {
    if (!b || world >= b->worlds.size()) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    HostWorld &w = b->worlds[world];
    if (w.resident == (resident != 0)) return BP_OK;
    bp_status s = b->download_state();            // every world's newest state, before the layout changes
    if (s != BP_OK) return s;
    w.resident = resident != 0;
    b->packed = false;
    return BP_OK;
}

bp_status bp_batch_world_footprint(const bp_batch *b, uint32_t world, uint64_t *bytes) {
    if (!b || world >= b->worlds.size() || !bytes) return BP_ERR_INVALID_ARGUMENT;
    *bytes = b->footprint(world);
    return BP_OK;
}

bp_status bp_batch_read_bubbles(bp_batch *b, uint32_t world, bp_bubble *out, uint32_t count) {
    if (!b || world >= b->worlds.size()) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_world(world);
    if (s != BP_OK) return s;
    const HostWorld &w = b->worlds[world];
    if (count != w.bubbles.size() || (count && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::copy(w.bubbles.begin(), w.bubbles.end(), out);
    return BP_OK;
}

bp_status bp_batch_read_all_bubbles(bp_batch *b, bp_bubble *out, uint32_t capacity, uint32_t *offsets, uint32_t offsets_count) {
    if (!b || !offsets || offsets_count != b->worlds.size() + 1) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_state();
    if (s != BP_OK) return s;
    uint32_t total = 0;
    for (size_t wi = 0; wi < b->worlds.size(); ++wi) { offsets[wi] = total; total += (uint32_t)b->worlds[wi].bubbles.size(); }
    offsets[b->worlds.size()] = total;
    if (!out) return BP_OK;
    if (capacity < total) return BP_ERR_CAPACITY_OVERFLOW;
    for (size_t wi = 0; wi < b->worlds.size(); ++wi)
        std::copy(b->worlds[wi].bubbles.begin(), b->worlds[wi].bubbles.end(), out + offsets[wi]);
    return BP_OK;
}

bp_status bp_batch_read_bodies(bp_batch *b, uint32_t world, bp_body *out, uint32_t count) {
    if (!b || world >= b->worlds.size()) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_state();
    if (s != BP_OK) return s;
    const HostWorld &hw = b->worlds[world];
    if (count != hw.bodies.size() || (count && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::copy(hw.bodies.begin(), hw.bodies.end(), out);
    return BP_OK;
}

float bp_batch_last_gpu_milliseconds(const bp_batch *b) { return b ? b->lastGpuMs : 0.0f; }

bp_status bp_batch_last_class_timing(const bp_batch *b, float *ms, uint32_t *worlds, uint32_t *threads, uint32_t *capacity) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < kClassCount; ++k) {
        if (ms) ms[k] = b->lastClassMs[k];
        if (worlds) worlds[k] = (uint32_t)b->classWorlds[k].size();
        if (threads) threads[k] = class_threads(k);
        if (capacity) capacity[k] = class_capacity(k);
    }
    return BP_OK;
}

uint32_t bp_batch_class_count(void) { return kClassCount; }

bp_status bp_batch_capture(bp_batch *b, uint32_t steps) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->capture(steps);
}

bp_status bp_batch_replay(bp_batch *b, bp_frame_info *infos) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(b->mutex, std::try_to_lock);
    if (!lock.owns_lock() || b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    b->inFlight = true;
    bp_status s = b->replay(infos);
    b->inFlight = false;
    return s;
}

bp_status bp_batch_audit(bp_batch *b, bp_audit *out, uint32_t count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->audit(out, count);
}

bp_status bp_batch_contacts(bp_batch *b, uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->contacts(world, out, capacity, out_count);
}

bp_status bp_batch_body_contacts(bp_batch *b, uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->contacts(world, out, capacity, out_count, true);
}

bp_status bp_batch_contours(bp_batch *b, uint32_t world, const bp_contour_config *config, uint32_t stride, float *points,
                            bp_bubble_contour *shapes, uint32_t bubble_count) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    return b->contours(world, config, stride, points, shapes, bubble_count);
}

bp_status bp_batch_set_world_class(bp_batch *b, uint32_t world, int32_t cls) {
    if (!b || world >= b->worlds.size() || cls < -1 || cls >= kClassCount) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_state();
    if (s != BP_OK) return s;
    if (cls >= 0 && !class_fits(cls, b->worlds[world])) return BP_ERR_INVALID_ARGUMENT;
    b->worlds[world].forcedClass = cls;
    b->packed = false;
    return BP_OK;
}

bp_status bp_batch_set_pooling(bp_batch *b, uint32_t enabled) {
    if (!b) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    b->pooling = enabled != 0;
    b->plan_pools();
    return BP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Plain world: a batch of one world (NativeMetalSolver).
// ---------------------------------------------------------------------------------------
struct bp_world {
    bp_batch batch;
    uint64_t publication = 1;          // counts published snapshots: full uploads and successful steps
    bool failNext = false;
    bool inFlight = false;
    std::vector<float> history;
    std::vector<uint32_t> rowOffsets, entries;
};

extern "C" {

bp_status bp_world_create(const bp_world_config *config, bp_world **out) {
    if (!config || !out) return BP_ERR_INVALID_ARGUMENT;
    bp_status s = device_available();
    if (s != BP_OK) return s;
    bp_world *w = new bp_world();
    HostWorld hw;
    hw.config = *config;
    hw.exportAdjacency = 1;
    w->batch.worlds.push_back(hw);
    w->batch.dropOnHold = false;       // the solver's frame keeps removed bubbles; the caller deletes them
    *out = w;
    return BP_OK;
}

void bp_world_destroy(bp_world *w) { delete w; }

static HostWorld &host_of(bp_world *w) { return w->batch.worlds[0]; }

static bp_status sync_host(bp_world *w) {
    bp_status s = w->batch.download_state();
    w->batch.packed = false;
    return s;
}

bp_status bp_world_set_config(bp_world *w, const bp_world_config *c) {
    if (!w || !c) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    host_of(w).config = *c;
    return BP_OK;
}

bp_status bp_world_set_bubbles(bp_world *w, const bp_bubble *bubbles, uint32_t count) {
    if (!w || (count && !bubbles)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    for (uint32_t i = 1; i < count; ++i)
        if (!(bubbles[i - 1].id < bubbles[i].id)) { set_error("bubbles must be sorted by strictly increasing id"); return BP_ERR_INVALID_SNAPSHOT; }
    host_of(w).bubbles.assign(bubbles, bubbles + count);
    ++w->publication;
    return BP_OK;
}

uint64_t bp_world_publication(const bp_world *w) { return w ? w->publication : 0u; }

bp_status bp_world_update_bubbles(bp_world *w, uint64_t base_publication, const uint32_t *indices, const bp_bubble *bubbles,
                                  uint32_t count) {
    if (!w || (count && (!indices || !bubbles))) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    if (base_publication != w->publication) {
        // NativeMetalSolver.swift:416-419: a delta must be taken against the last published snapshot.
        set_error("the delta is not based on the last published snapshot");
        return BP_ERR_INVALID_SNAPSHOT;
    }
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    HostWorld &hw = host_of(w);
    for (uint32_t k = 0; k < count; ++k) {
        if (indices[k] >= hw.bubbles.size()) return BP_ERR_INVALID_SNAPSHOT;
        if (bubbles[k].id != hw.bubbles[indices[k]].id) { set_error("delta changes the id of a bubble"); return BP_ERR_INVALID_SNAPSHOT; }
    }
    for (uint32_t k = 0; k < count; ++k) hw.bubbles[indices[k]] = bubbles[k];
    return BP_OK;
}

bp_status bp_world_set_segments(bp_world *w, const bp_segment *segments, uint32_t count) {
    if (!w || (count && !segments)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    host_of(w).segments.assign(segments, segments + count);
    return BP_OK;
}

bp_status bp_world_set_bodies(bp_world *w, const bp_body *bodies, uint32_t count, const float *vertices, uint32_t vertex_count,
                              const uint32_t *triangles, uint32_t triangle_count) {
    if (!w || (count && !bodies) || (vertex_count && !vertices) || (triangle_count && !triangles)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    HostWorld &hw = host_of(w);
    hw.bodies.assign(bodies, bodies + count);
    hw.vertices.assign(vertices, vertices + 2 * (size_t)vertex_count);
    hw.triangles.assign(triangles, triangles + 4 * (size_t)triangle_count);
    return BP_OK;
}

bp_status bp_world_set_forces(bp_world *w, const bp_motor *motors, uint32_t motor_count, const bp_band *bands, uint32_t band_count) {
    if (!w || (motor_count && !motors) || (band_count && !bands)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    HostWorld &hw = host_of(w);
    hw.motors.assign(motors, motors + motor_count);
    hw.bands.assign(bands, bands + band_count);
    return BP_OK;
}

bp_status bp_world_set_fields(bp_world *w, const bp_field *fields, uint32_t field_count, const float *vertices,
                              uint32_t vertex_count, const int64_t *filter_ids, uint32_t filter_count) {
    if (!w || (field_count && !fields) || (vertex_count && !vertices) || (filter_count && !filter_ids)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = sync_host(w);
    if (s != BP_OK) return s;
    HostWorld &hw = host_of(w);
    hw.fields.assign(fields, fields + field_count);
    hw.fieldVertices.assign(vertices, vertices + 2 * (size_t)vertex_count);
    hw.fieldIds.assign(filter_ids, filter_ids + filter_count);
    return BP_OK;
}

bp_status bp_world_step(bp_world *w, int32_t scenario_step, bp_frame_info *out) {
    (void)scenario_step;
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    if (w->failNext) {
        w->failNext = false;
        set_error("command executor failure requested by the test switch");
        if (out) { std::memset(out, 0, sizeof *out); out->status = BP_ERR_COMMAND_FAILED; }
        return BP_ERR_COMMAND_FAILED;
    }
    w->inFlight = true;
    bp_frame_info info{};
    bp_status s = BP_OK;
    for (int attempt = 0; attempt < 4; ++attempt) {
        s = w->batch.step(1, &info);
        if (s != BP_OK) break;
        if (info.status != BP_ERR_CAPACITY_OVERFLOW) break;
        // The adjacency did not fit: nothing was published; grow and run the step again.
        HostWorld &hw = host_of(w);
        s = w->batch.download_state();
        if (s != BP_OK) break;
        hw.adjPerBubble *= 4;
        w->batch.packed = false;
    }
    w->inFlight = false;
    if (s != BP_OK) {
        if (out) { std::memset(out, 0, sizeof *out); out->status = s; }
        return s;
    }
    // A failed world keeps its published state and may be stepped again after new input.
    if (info.status != BP_OK) {
        bpd::WorldDesc d;
        if (hipMemcpy(&d, w->batch.descs, sizeof d, hipMemcpyDeviceToHost) == hipSuccess && d.failed) {
            d.failed = 0;
            (void)hipMemcpy(w->batch.descs, &d, sizeof d, hipMemcpyHostToDevice);
        }
    }
    if (info.status == BP_OK) ++w->publication;
    if (out) *out = info;
    return info.status;
}

bp_status bp_world_region_results(bp_world *w, int64_t *removed, uint32_t removed_capacity, uint32_t *out_removed,
                                  int64_t *wrapped, uint32_t wrapped_capacity, uint32_t *out_wrapped) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    return w->batch.region_results(0, removed, removed_capacity, out_removed, wrapped, wrapped_capacity, out_wrapped);
}

bp_status bp_world_remove_bubbles(bp_world *w, const int64_t *ids, uint32_t count) {
    if (!w || (count && !ids)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = w->batch.remove_bubbles(0, ids, count);
    if (s == BP_OK) ++w->publication;
    return s;
}

bp_status bp_world_read_bubbles(bp_world *w, bp_bubble *out, uint32_t count) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    bp_status s = w->batch.download_state();
    if (s != BP_OK) return s;
    const HostWorld &hw = host_of(w);
    if (count != hw.bubbles.size() || (count && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::copy(hw.bubbles.begin(), hw.bubbles.end(), out);
    return BP_OK;
}

bp_status bp_world_read_bodies(bp_world *w, bp_body *out, uint32_t count) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    bp_status s = w->batch.download_state();
    if (s != BP_OK) return s;
    const HostWorld &hw = host_of(w);
    if (count != hw.bodies.size() || (count && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::copy(hw.bodies.begin(), hw.bodies.end(), out);
    return BP_OK;
}

uint32_t bp_world_bubble_count(const bp_world *w) { return w ? (uint32_t)w->batch.worlds[0].bubbles.size() : 0u; }

bp_status bp_world_read_adjacency(bp_world *w, uint32_t *row_offsets, uint32_t offsets_count, uint32_t *entries,
                                  uint32_t entries_capacity, uint32_t *out_entry_count) {
    if (!w || !w->batch.packed) return BP_ERR_INVALID_ARGUMENT;
    const bpd::WorldDesc &d = w->batch.hostDescs[0];
    uint32_t n = d.bubbleCount;
    if (!row_offsets || offsets_count != n + 1) return BP_ERR_INVALID_ARGUMENT;
    std::vector<uint32_t> start(n), len(n);
    if (n) {
        BP_HIP(hipMemcpy(start.data(), w->batch.arena.rowStart + d.bubbleOffset, n * 4, hipMemcpyDeviceToHost));
        BP_HIP(hipMemcpy(len.data(), w->batch.arena.rowLen + d.bubbleOffset, n * 4, hipMemcpyDeviceToHost));
    }
    uint32_t total = 0;
    for (uint32_t i = 0; i < n; ++i) { row_offsets[i] = total; total += len[i]; }
    row_offsets[n] = total;
    if (out_entry_count) *out_entry_count = total;
    if (entries && total) {
        if (entries_capacity < total) return BP_ERR_CAPACITY_OVERFLOW;
        uint32_t cap = d.adjCapacity;
        std::vector<uint32_t> all(cap);
        BP_HIP(hipMemcpy(all.data(), w->batch.arena.adj + d.adjOffset, (size_t)cap * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t k = 0; k < len[i]; ++k) entries[row_offsets[i] + k] = all[start[i] + k];
    }
    return BP_OK;
}

bp_status bp_world_read_step_history(bp_world *w, float *out, uint32_t capacity, uint32_t *out_count) {
    if (!w || !w->batch.packed || !out) return BP_ERR_INVALID_ARGUMENT;
    uint32_t count = std::min<uint32_t>(capacity, bpd::kFrameLimit);
    const bpd::WorldDesc &d = w->batch.hostDescs[0];
    BP_HIP(hipMemcpy(out, w->batch.arena.history + d.historyOffset, count * sizeof(float), hipMemcpyDeviceToHost));
    if (out_count) *out_count = w->batch.hostStats.empty() ? 0u : std::min(w->batch.hostStats[0].substeps, count);
    return BP_OK;
}

bp_status bp_world_audit(bp_world *w, bp_audit *out) {
    if (!w || !out) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    return w->batch.audit(out, 1);
}

bp_status bp_world_contacts(bp_world *w, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!w || !out_count) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    return w->batch.contacts(0, out, capacity, out_count);
}

bp_status bp_world_contour_stride(bp_world *w, const bp_contour_config *config, uint32_t *out_stride) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    return w->batch.contour_stride(0, config, out_stride);
}

bp_status bp_world_contours(bp_world *w, const bp_contour_config *config, uint32_t stride, float *points,
                            bp_bubble_contour *shapes, uint32_t bubble_count) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    return w->batch.contours(0, config, stride, points, shapes, bubble_count);
}

bp_status bp_world_body_contacts(bp_world *w, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!w || !out_count) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    return w->batch.contacts(0, out, capacity, out_count, true);
}

// bp_world_contacts without any allocation: the slot offsets and records live in the library's
// fixed per-GPU scratch array (bp_contacts.hip), the records are read back into the caller's
// array and widened to bp_contact there, in place.
bp_status bp_world_contacts_noalloc(bp_world *w, bp_contact *out, uint32_t capacity, uint32_t *out_count) {
    if (!w || !out_count || (capacity && !out)) return BP_ERR_INVALID_ARGUMENT;
    if (w->inFlight) return BP_ERR_FRAME_IN_FLIGHT;
    bp_batch &b = w->batch;
    if (!b.packed || !b.deviceValid || !b.descs) {
        set_error("the world has no state on the device: upload and step it before reading contacts");
        return BP_ERR_UNAVAILABLE;
    }
    void *base = nullptr;
    uint32_t maxBubbles = 0, maxRecords = 0;
    hipError_t e = bp_contacts_scratch(&base, &maxBubbles, &maxRecords);
    if (e != hipSuccess) { set_error(std::string("scratch address: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_SETUP_FAILED; }
    bpd::WorldDesc d;
    BP_HIP(hipMemcpy(&d, b.descs, sizeof d, hipMemcpyDeviceToHost));
    if (!d.exportAdjacency) { set_error("contact records read the rows of the last grid build"); return BP_ERR_UNAVAILABLE; }
    if (d.bubbleCount > maxBubbles || capacity > maxRecords) {
        set_error("bp_world_contacts_noalloc holds worlds of at most 32768 bubbles and 65536 records per call");
        return BP_ERR_CAPACITY_OVERFLOW;
    }
    *out_count = 0;
    if (d.bubbleCount == 0) return BP_OK;
    uint32_t *offsets = static_cast<uint32_t *>(base);
    char *records = static_cast<char *>(base) + (size_t)2 * maxBubbles * sizeof(uint32_t);
    uint32_t *total = reinterpret_cast<uint32_t *>(records + (size_t)maxRecords * sizeof(bpd::ContactRecord));
    hipStream_t stream = b.streams[0];
    e = bp_launch_contacts(b.arenaDev, b.descs, 0, offsets, records, capacity, total, stream);
    uint32_t kept = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&kept, total, sizeof kept, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    const uint32_t n = std::min(kept, capacity);
    // The device records (56 bytes) land at the front of the caller's array (64-byte records)...
    if (e == hipSuccess && n) e = hipMemcpy(out, records, (size_t)n * sizeof(bpd::ContactRecord), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error(std::string("contacts: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    // ...and are widened in place from the last: record k moves from 56 k to 64 k, so it is read
    // whole before any write reaches it, and only records already moved are overwritten.
    const char *raw = reinterpret_cast<const char *>(out);
    for (uint32_t k = n; k-- > 0;) {
        bpd::ContactRecord r;
        std::memcpy(&r, raw + (size_t)k * sizeof(bpd::ContactRecord), sizeof r);
        bp_contact c;
        uint64_t h = 14695981039346656037ull;          // NativeMetalSolver.swift:1047-1052
        h = (h ^ (uint64_t)r.kind) * 1099511628211ull;
        h = (h ^ (uint64_t)r.first) * 1099511628211ull;
        h = (h ^ (uint64_t)r.second) * 1099511628211ull;
        c.id = h;
        c.first = r.first; c.second = r.second;
        c.kind = r.kind; c.feature = r.feature;
        c.normal[0] = r.normal[0]; c.normal[1] = r.normal[1];
        c.point[0] = r.point[0]; c.point[1] = r.point[1];
        c.penetration = r.penetration;
        c.allowed_side = r.allowedSide;
        c.impulse = r.impulse;
        c.index_b = r.indexB;
        std::memcpy(reinterpret_cast<char *>(out) + (size_t)k * sizeof(bp_contact), &c, sizeof c);
    }
    *out_count = kept;
    return BP_OK;
}

bp_status bp_world_debug_fail_next_command(bp_world *w) {
    if (!w) return BP_ERR_INVALID_ARGUMENT;
    w->failNext = true;
    return BP_OK;
}

}  // extern "C"
