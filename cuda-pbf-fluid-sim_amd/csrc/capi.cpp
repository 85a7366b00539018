// capi.cpp — C handle API over sim::Simulator.
#include <cstring>
#include <new>

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
    // step(p) applies the parameters and advances once; here only apply them.
    sim::SimParams p = *params;
    if (std::memcmp(&p, &sim->s.params(), sizeof(p)) == 0) return 0;
    if (!sim->s.step(p)) return -1;
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
