// capi.cpp — C handle API over sim::Simulator.
#include <cstring>
#include <new>
#include <vector>

#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_sim.h"
#include "sim_host.h"

struct PbfSim {
    sim::Simulator s;
};

extern "C" PbfSim* pbfSimCreate(const sim::SimParams* params, int gridKind, int tableLog2, int path, int useGraph) {
    if (!params || gridKind < 0 || gridKind > 2 || path < 0 || path > 1) {
        pbfx::set_error(PBFX_E_INVALID_ARGUMENT, "pbfSimCreate: bad arguments");
        return nullptr;
    }
    PbfSim* h = new (std::nothrow) PbfSim;
    if (!h) return nullptr;
    sim::SimOptions opt;
    opt.grid = (sim::GridKind)gridKind;
    opt.path = (sim::PathKind)path;
    opt.tableLog2 = tableLog2 > 0 ? (uint32_t)tableLog2 : 22u;
    opt.useGraph = useGraph != 0;
    if (!h->s.initialize(*params, opt)) {
        delete h;
        return nullptr;
    }
    return h;
}

extern "C" void pbfSimDestroy(PbfSim* sim) { delete sim; }

extern "C" int pbfSimSeedLattice(PbfSim* sim, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float ox,
                                 float oy, float oz, float spacing, float jitterAmp, uint64_t seed) {
    if (!sim) return -1;
    return sim->s.seedBoxLattice(first, nx, ny, nz, make_float3(ox, oy, oz), spacing, jitterAmp, seed) ? 0 : -1;
}

extern "C" int pbfSimSeedCubeMix(PbfSim* sim, uint32_t first, uint32_t groupCount, const float* centers,
                                 uint32_t edgeParticles, float spacing, float jitterAmp, uint64_t seed) {
    if (!sim || (groupCount && !centers)) return -1;
    std::vector<float3> c(groupCount);
    for (uint32_t g = 0; g < groupCount; ++g) c[g] = make_float3(centers[3 * g], centers[3 * g + 1], centers[3 * g + 2]);
    return sim->s.seedCubeMix(first, groupCount, c.data(), edgeParticles, spacing, jitterAmp, seed) ? 0 : -1;
}

extern "C" int pbfSimBindExternalPositions(PbfSim* sim, float4* posA, float4* posB, uint32_t capacity) {
    if (!sim) return -1;
    return sim->s.bindExternalPositions(posA, posB, capacity) ? 0 : -1;
}

extern "C" int pbfSimBindExternalVelocity(PbfSim* sim, void* dst, size_t bytes, uint32_t strideBytes) {
    if (!sim) return -1;
    return sim->s.bindExternalVelocity(dst, bytes, strideBytes) ? 0 : -1;
}

extern "C" int pbfSimPublishVelocity(PbfSim* sim, uint32_t count) {
    if (!sim) return -1;
    return sim->s.publishVelocity(count) ? 0 : -1;
}

extern "C" int pbfSimSetTuning(PbfSim* sim, const PbfSimTuning* t) {
    if (!sim || !t) return -1;
    const sim::Simulator::Tuning k{t->candidateMask != 0, t->streamRatio, t->chunkBits != 0, t->sortItems,
                              t->sortDigitBits,      t->walkBlock,  t->laneOrder != 0};
    return sim->s.setTuning(k) ? 0 : -1;
}

extern "C" int pbfSimGetTuning(PbfSim* sim, PbfSimTuning* t) {
    if (!sim || !t) return -1;
    const sim::Simulator::Tuning k = sim->s.tuning();
    t->candidateMask = k.candidateMask ? 1 : 0;
    t->streamRatio = k.streamRatio;
    t->chunkBits = k.chunkBits ? 1 : 0;
    t->sortItems = k.sortItems;
    t->sortDigitBits = k.sortDigitBits;
    t->walkBlock = k.walkBlock;
    t->laneOrder = k.laneOrder ? 1 : 0;
    return 0;
}

extern "C" int pbfSimSetState(PbfSim* sim, const float4* pos, const float4* vel, uint32_t N, int fromDevice) {
    if (!sim) return -1;
    return sim->s.setState(pos, vel, N, fromDevice != 0) ? 0 : -1;
}

extern "C" int pbfSimGetState(PbfSim* sim, float4* pos, float4* vel, int toDevice) {
    if (!sim) return -1;
    return sim->s.getState(pos, vel, toDevice != 0) ? 0 : -1;
}

extern "C" int pbfSimStep(PbfSim* sim, int steps) {
    if (!sim || steps < 0) return -1;
    return sim->s.stepMany(steps) ? 0 : -1;
}

extern "C" int pbfSimSetParams(PbfSim* sim, const sim::SimParams* params) {
    if (!sim || !params) return -1;
    return sim->s.setParams(*params) ? 0 : -1;
}

extern "C" int pbfSimGetParams(PbfSim* sim, sim::SimParams* out) {
    if (!sim || !out) return -1;
    *out = sim->s.params();
    return 0;
}

extern "C" int pbfSimSynchronize(PbfSim* sim) {
    if (!sim) return -1;
    return hipStreamSynchronize(sim->s.stream()) == hipSuccess ? 0 : -1;
}

extern "C" hipStream_t pbfSimStream(PbfSim* sim) { return sim ? sim->s.stream() : nullptr; }
extern "C" const float4* pbfSimDevicePositions(PbfSim* sim) { return sim ? sim->s.devicePositions() : nullptr; }
extern "C" const float4* pbfSimDeviceVelocities(PbfSim* sim) { return sim ? sim->s.deviceVel() : nullptr; }
extern "C" const uint32_t* pbfSimDeviceSortedKeys(PbfSim* sim) {
    return sim ? sim->s.grid().d_cellKeys_sorted : nullptr;
}
extern "C" const uint32_t* pbfSimDeviceSortedIndices(PbfSim* sim) {
    return sim ? sim->s.grid().d_indices_sorted : nullptr;
}

extern "C" int pbfSimProfilePhases(PbfSim* sim, float* ms, int maxPhases, char* names, int namesBytes) {
    if (!sim) return -1;
    std::vector<std::pair<std::string, float>> out;
    if (!sim->s.profilePhases(out)) return -1;
    int pos = 0;
    for (int i = 0; i < (int)out.size() && i < maxPhases; ++i) {
        if (ms) ms[i] = out[i].second;
        if (names) {
            const std::string& n = out[i].first;
            if (pos + (int)n.size() + 1 <= namesBytes) {
                std::memcpy(names + pos, n.c_str(), n.size() + 1);
                pos += (int)n.size() + 1;
            }
        }
    }
    return (int)out.size();
}

extern "C" int pbfSimDensityStats(PbfSim* sim, double* out5) {
    if (!sim || !out5) return -1;
    return sim->s.densityStats(out5) ? 0 : -1;
}

extern "C" int pbfSimHashOverflow(PbfSim* sim) { return (sim && sim->s.hashOverflowed()) ? 1 : 0; }

extern "C" void pbfDebugEchoDeviceParams(sim::DeviceParams dp, sim::DeviceParams* out) {
    if (out) *out = dp;
}
extern "C" void pbfDebugEchoGridBounds(sim::GridBounds g, sim::GridBounds* out) {
    if (out) *out = g;
}
extern "C" void pbfDebugEchoFloat3(float3 v, float* out3) {
    if (out3) {
        out3[0] = v.x;
        out3[1] = v.y;
        out3[2] = v.z;
    }
}

static_assert(sizeof(PbfSceneDesc) == 48, "PbfSceneDesc layout");

extern "C" PbfSim* pbfEnsembleCreate(const sim::SimParams* base, const PbfSceneDesc* scenes, uint32_t count,
                                     int gridKind, int tableLog2, int useGraph) {
    if (!base || !scenes || count == 0 || gridKind < 0 || gridKind > 2) {
        pbfx::set_error(PBFX_E_INVALID_ARGUMENT, "pbfEnsembleCreate: bad arguments");
        return nullptr;
    }
    std::vector<sim::SceneDesc> d(count);
    for (uint32_t i = 0; i < count; ++i) {
        const PbfSceneDesc& c = scenes[i];
        d[i].boxMin = make_float3(c.boxMin[0], c.boxMin[1], c.boxMin[2]);
        d[i].boxMax = make_float3(c.boxMax[0], c.boxMax[1], c.boxMax[2]);
        d[i].gravity = make_float3(c.gravity[0], c.gravity[1], c.gravity[2]);
        d[i].xsph_c = c.xsph_c;
        d[i].restitution = c.restitution;
        d[i].numParticles = c.numParticles;
    }
    PbfSim* h = new (std::nothrow) PbfSim;
    if (!h) return nullptr;
    sim::SimOptions opt;
    opt.grid = (sim::GridKind)gridKind;
    opt.path = sim::PathKind::Reordered;
    opt.tableLog2 = tableLog2 > 0 ? (uint32_t)tableLog2 : 22u;
    opt.useGraph = useGraph != 0;
    if (!h->s.initializeEnsemble(*base, d, opt)) {
        delete h;
        return nullptr;
    }
    return h;
}

extern "C" int pbfEnsembleSeedLattice(PbfSim* sim, uint32_t scene, uint32_t nx, uint32_t ny, uint32_t nz, float ox,
                                      float oy, float oz, float spacing, float jitterAmp, uint64_t seed) {
    if (!sim) return -1;
    return sim->s.seedSceneLattice(scene, nx, ny, nz, make_float3(ox, oy, oz), spacing, jitterAmp, seed) ? 0 : -1;
}

extern "C" int pbfEnsembleGetScene(PbfSim* sim, uint32_t scene, float4* pos, float4* vel) {
    if (!sim) return -1;
    return sim->s.getSceneState(scene, pos, vel) ? 0 : -1;
}

extern "C" int pbfEnsembleSceneInfo(PbfSim* sim, uint32_t scene, uint32_t* first, uint32_t* count, int* dims3) {
    if (!sim) return -1;
    const pbfx::SceneParams* sp = sim->s.sceneParams(scene);
    if (!sp) return -1;
    if (first) *first = sp->particleBase;
    if (count) *count = sp->numParticles;
    if (dims3) {
        dims3[0] = sp->grid.dim.x;
        dims3[1] = sp->grid.dim.y;
        dims3[2] = sp->grid.dim.z;
    }
    return 0;
}

extern "C" int pbfSimAdjointInit(PbfSim* sim, int tapeSlots) {
    if (!sim) return -1;
    return sim->s.adjointInit(tapeSlots) ? 0 : -1;
}

extern "C" int pbfSimStepRecord(PbfSim* sim, int slot) {
    if (!sim) return -1;
    return sim->s.stepRecord(slot) ? 0 : -1;
}

extern "C" int pbfSimBackwardStep(PbfSim* sim, int slot, const float4* gXin, const float4* gVin, float4* gXout,
                                  float4* gVout, double* gcAccum) {
    if (!sim || !gXin || !gVin || !gXout || !gVout || !gcAccum) return -1;
    return sim->s.backwardStep(slot, gXin, gVin, gXout, gVout, gcAccum) ? 0 : -1;
}

extern "C" int pbfSimBackwardStepScaled(PbfSim* sim, int slot, const float4* gXin, const float4* gVin,
                                        float4* gXout, float4* gVout, double* gcAccum, int32_t* log2Scale) {
    if (!sim || !gXin || !gVin || !gXout || !gVout || !gcAccum || !log2Scale) return -1;
    return sim->s.backwardStep(slot, gXin, gVin, gXout, gVout, gcAccum, log2Scale) ? 0 : -1;
}
