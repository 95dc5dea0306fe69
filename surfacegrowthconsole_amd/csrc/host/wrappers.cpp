// wrappers.cpp - the reference's extern "C" wrapper API (SurfaceGrowth.h:217-235) on top of the
// sgk_* engine.  DoComputationsW keeps the reference step loop's observable behaviour
// (SurfaceGrowthConsole/SurfaceGrowth.cu:2013-2529): same outputs, same error convention.  The
// difference is where the work runs: whole chunks of steps are device resident and the host
// consumes the per-step records afterwards, synchronising only where an output needs device data
// (pdb snapshots, RDF prints, backups, the end of the run).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_internal.h"
#include "sgk.h"

namespace {

SimParams g_params;  // copy made by SetParametersW (the reference copies to __constant__ memory)

char *fail(char *szPdbPath, const char *what, const char *detail) {
    std::string s = std::string(what) + (detail ? detail : "");
    std::strncpy(szPdbPath, s.c_str(), MAX_PATH - 1);
    szPdbPath[MAX_PATH - 1] = 0;
    return szPdbPath;
}

long long next_multiple(long long s, long long k) { return (s / k + 1) * k; }

// Restart from backup (row 22, bStartBckup).  The reference parses the row (SurfaceGrowth.cpp:775,
// stored at :378) and writes backups (SurfaceGrowth.cu:2457-2480, alternately to szBckup0/1) but
// never reads one; here the run continues from the newer of the two files (the larger stepCount in
// its SimParams).  Restored: positions, velocities, accelerations and the run state the backup's
// SimParams carries (step, displacement, deposition counter, averaging accumulators, shear, elapsed
// time).  The first step rebuilds the neighbour list, which a backup does not hold; where the
// continuous run rebuilds at that step too (a list rebuilt every k steps, backups every multiple of
// k), the restarted run is the continuous one, bit for bit.  The RDF histogram and the diffusion buffers are not in a backup: both
// start empty.  The input's other fields (step limit, intervals, outputs) stay those of this run.
// Returns 0, or the message.
const char *restore_backup(SimParams &p, sg_float4 *hr, sg_float3 *hv, sg_float3 *ha) {
    const char *paths[2] = {p.szBckup0, p.szBckup1};
    int best = -1;
    SimParams b;
    for (int k = 0; k < 2; ++k) {
        FILE *f = std::fopen(paths[k], "rb");
        if (!f) continue;
        SimParams h;
        const bool ok = std::fread(&h, sizeof(SimParams), 1, f) == 1;
        std::fclose(f);
        if (ok && (best < 0 || h.stepCount > b.stepCount)) {
            best = k;
            b = h;
        }
    }
    if (best < 0) return "Cannot open backup file!";
    if (b.nMol != p.nMol || b.nMolMe != p.nMolMe || b.iRegime != p.iRegime || b.iMaterial != p.iMaterial ||
        b.region.x != p.region.x || b.region.y != p.region.y || b.region.z != p.region.z ||
        b.cells.x != p.cells.x || b.cells.y != p.cells.y || b.cells.z != p.cells.z)
        return "The backup file does not match the input!";
    if (!sgReadBackup(paths[best], &b, hr, hv, ha, p.nMol)) return "Cannot read backup file!";
    p.stepCount = b.stepCount;
    p.nebrNow = 1;  // the list is not in a backup: the first step builds it
    p.dispHi = b.dispHi;
    p.nMolDeposited = b.nMolDeposited;
    p.vSum = b.vSum;
    p.kinEnergy = b.kinEnergy;
    p.totEnergy = b.totEnergy;
    p.potEnergy = b.potEnergy;
    p.oneStep = b.oneStep;
    p.centerOfMass = b.centerOfMass;
    p.cmVel = b.cmVel;
    p.frictForce = b.frictForce;
    p.particleSize = b.particleSize;
    p.shear = b.shear;
    p.totalShear = b.totalShear;
    p.totalTime = b.totalTime;
    p.countRdf = 0;
    p.countDiffuseAv = 0;
    p.moreCycles = p.stepCount < p.stepLimit ? 1 : 0;
    return 0;
}

}  // namespace

extern "C" {

void CudaInitW(int argc, char **argv) {
    (void)argc;
    (void)argv;
    int n = sgk_device_count();
    if (n < 0) {
        std::printf("hipGetDeviceCount failed\n");
        std::exit(EXIT_FAILURE);
    }
    if (n == 0)
        std::printf("There are no available device(s) that support HIP\n");
    else
        std::printf("Detected %d HIP capable device(s)\n", n);
}

int sgDeviceCount(void) {
    const int n = sgk_device_count();
    return n > 0 ? n : 0;
}

unsigned int sgMaxCellGrid(void) { return (unsigned int)((2147483647u) / SG_BLOCK_SIZE); }

void SetParametersW(SimParams *hostParams) {
    if (!hostParams) {
        std::printf("Memory error while copying parameters\n");
        std::exit(EXIT_FAILURE);
    }
    g_params = *hostParams;
}

// InitCoordsW (SurfaceGrowth.cu:2532-2609).  `dr` is ignored as in the reference.
const char *InitCoordsW(sg_float4 *dr, sg_float4 *hr, SimParams *hp) {
    (void)dr;
    const SgOptions &o = sg::options();
    if (!hr || !hp) return "null argument";
    if (hp->iRegime == 0) {
        sg::gen_fcc_bulk(*hp, hr);
    } else if (hp->iRegime == 1) {
        sg::gen_graphene(*hp, hr);
        sg::gen_deposit_random(*hp, hr);
    } else if (hp->iRegime == 2) {
        sg::gen_graphene(*hp, hr);
        if (hp->nMolMe != 0) {
            const double gap = o.gapAngstrom / hp->lengthU;
            if (o.geometry == SG_GEOM_SPHERE) {
                if (sg::gen_sphere(*hp, gap, hr)) return "sphere generation failed";
            } else if (o.geometry == SG_GEOM_ICOSAHEDRON) {
                if (sg::gen_icosahedron(*hp, gap, hr))
                    return "icosahedron generation failed: nMolMe must be a Mackay magic number (13, 55, 147, 309, 561, ...)";
            } else {
                sg::gen_slab(*hp, hr);
            }
        }
    } else {
        return "unknown regime";
    }
    return 0;
}

char *DoComputationsW(sg_float4 *hr, sg_float3 *hv, sg_float3 *ha, SimParams *hparams, FILE *fResults,
                      char *szPdbPath) {
    SimParams &p = *hparams;
    const SgOptions &o = sg::options();
    if (p.bStartBckup) {
        if (!hr || !hv || !ha) return fail(szPdbPath, "Cannot restart from backup: ", "null state arrays");
        if (const char *e = restore_backup(p, hr, hv, ha)) return fail(szPdbPath, e, nullptr);
    }
    std::unique_ptr<sg::Diffusion> diff;
    FILE *fileDiffuse = nullptr;
    if (p.iRegime == 2) {
        diff.reset(new sg::Diffusion(p));
        if ((fileDiffuse = std::fopen(p.szDiffusePath, "w")) == nullptr)
            return fail(szPdbPath, "Cannot open diffuse file!", nullptr);
    }
    const int chunkMax = std::max(1, std::min(o.stepsPerSync, 4096));
    sgk_config cfg;
    sgk_default_config(&cfg);
    cfg.precision = o.precision;
    cfg.rebuild_every = o.rebuildEvery;
    cfg.use_graphs = o.useGraphs;
    cfg.records_capacity = chunkMax;
    cfg.metal_thermostat_all = o.metalThermostatAll;
    sgk_ctx *ctx = nullptr;
    if (sgk_create(&p, &cfg, &ctx) != SGK_OK) {
        if (fileDiffuse) std::fclose(fileDiffuse);
        return fail(szPdbPath, "Problems with memory allocation! Exception: ", sgk_last_error(nullptr));
    }
    auto bail = [&](const char *what) {
        std::string msg = sgk_last_error(ctx);
        sgk_destroy(ctx);
        if (fileDiffuse) std::fclose(fileDiffuse);
        return fail(szPdbPath, what, msg.c_str());
    };
    if (sgk_upload(ctx, hr, hv, ha) != SGK_OK) return bail("Problems with memory allocation! Exception: ");
    if (sgk_set_run_state(ctx, p.stepCount, p.nebrNow, p.dispHi) != SGK_OK ||
        sgk_set_deposited(ctx, p.nMolDeposited) != SGK_OK)
        return bail("Problems with memory allocation! Exception: ");
    // a restarted shear run continues with its shear force; the last step's centre of mass and its
    // velocity are the backup's FP32 values
    if (p.bStartBckup && sgk_set_shear_state(ctx, p.shear, p.centerOfMass.val, p.cmVel.val) != SGK_OK)
        return bail("Problems with memory allocation! Exception: ");
    if (p.bPdb != 0) sg::create_pdb_file(szPdbPath, p, hr);

    std::vector<sgk_record> recs(chunkMax);
    std::vector<unsigned> hist(p.sizeHistRdf > 0 ? p.sizeHistRdf : 1);
    const int rdfEvery = (p.bRdf != 0 && p.stepRdf > 0) ? p.stepRdf : 0;
    int iBckup = 1;
    while (p.moreCycles) {
        const long long s0 = p.stepCount;
        long long end = std::min<long long>(p.stepLimit, s0 + chunkMax);
        if (end <= s0) end = s0 + 1;  // the reference always runs at least one step
        if (p.bPdb != 0 && p.stepPdb > 0) end = std::min(end, next_multiple(s0, p.stepPdb));
        if (p.bBckup && p.stepBckup > 0) end = std::min(end, next_multiple(s0, p.stepBckup));
        if (rdfEvery > 0) {
            const long long need = std::max(1, p.limitRdf - p.countRdf);
            end = std::min(end, next_multiple(s0, rdfEvery) + (need - 1) * rdfEvery);
        }
        const int n = (int)(end - s0);
        if (sgk_run(ctx, n, rdfEvery, recs.data()) != SGK_OK) return bail("Problems with evaluation of properties! Exception: ");
        const float msPerStep = (float)(sgk_last_run_ms(ctx) / n);
        for (int k = 0; k < n; ++k) {
            const sgk_record &r = recs[k];
            p.stepCount = r.step;
            p.nMolDeposited = r.nMolDeposited;
            if (p.nMolMe != 0) {
                p.centerOfMass.val = (float)r.comSum[0] / p.nMolMe;
                p.cmVel.val = (float)r.cmVelSum[0] / p.nMolMe;
                p.frictForce.val = (float)r.frict;
                p.particleSize.x = (float)r.bboxMax[0] - (float)r.bboxMin[0];
                p.particleSize.y = (float)r.bboxMax[1] - (float)r.bboxMin[1];
                p.particleSize.z = (float)r.bboxMax[2] - (float)r.bboxMin[2];
                if (p.iRegime == 2 && p.stepCount > p.stepEquil &&
                    (p.stepCount - p.stepEquil) % p.stepDiffuse == 0) {
                    sg_float3 com = {(float)r.comSum[0] / p.nMolMe, (float)r.comSum[1] / p.nMolMe,
                                     (float)r.comSum[2] / p.nMolMe};
                    diff->eval(p, com, fileDiffuse);
                }
            } else {
                p.centerOfMass.val = p.cmVel.val = p.frictForce.val = 0.f;
                p.particleSize = {0.f, 0.f, 0.f};
            }
            p.dispHi = (float)r.dispHi;
            p.nebrNow = r.nebrNow;
            p.kinEnergy.val = (float)r.kinEnergy;
            p.potEnergy.val = (float)r.potEnergy;
            p.totEnergy.val = (float)r.totEnergy;
            p.vSum = {(float)r.vSum[0], (float)r.vSum[1], (float)r.vSum[2]};
            p.shear = (float)r.shear;
            if (p.iRegime == 2 && p.nMolMe != 0 && p.stepCount > p.stepEquil + p.stepCool)
                p.totalShear = (float)r.shearCount * p.shear;
            if (rdfEvery > 0 && p.stepCount % rdfEvery == 0) {
                ++p.countRdf;
                if (p.countRdf == p.limitRdf) {  // only at the last step of a chunk (see `end`)
                    if (sgk_rdf_get(ctx, hist.data()) != SGK_OK) return bail("Problems with rdf! Exception: ");
                    sg::print_rdf(p, hist.data());
                    sgk_rdf_clear(ctx);
                    p.countRdf = 0;
                }
            }
            p.oneStep.val = msPerStep;
            sg::accum_props(1, p);
            if (p.stepCount % p.stepAvg == 0) {
                sg::accum_props(2, p);
                if (p.bResult != 0 && fResults) sg::print_summary(fResults, p);
                sg::accum_props(0, p);
            }
            const bool pdbNow = p.stepPdb > 0 && p.stepCount % p.stepPdb == 0 && p.bPdb != 0;
            const bool bckNow = p.bBckup && p.stepBckup > 0 && p.stepCount % p.stepBckup == 0;
            if (pdbNow || bckNow) {  // only at the last step of a chunk
                if (sgk_download(ctx, hr, bckNow ? hv : nullptr, bckNow ? ha : nullptr) != SGK_OK)
                    return bail("Problems with copying coordinates! Exception: ");
                if (pdbNow) sg::create_pdb_file(szPdbPath, p, hr);
                if (bckNow) {
                    sgWriteBackup(iBckup ? p.szBckup0 : p.szBckup1, &p, hr, hv, ha);
                    iBckup = !iBckup;
                }
            }
            if (p.stepCount >= p.stepLimit) p.moreCycles = 0;
            p.totalTime += msPerStep * 0.001f;
        }
    }
    if (p.bResult != 0 && fResults) std::fprintf(fResults, "\nDuration of the simulation = %f s", p.totalTime);
    // Final state back to the caller's arrays (the reference leaves them at the last snapshot).
    sgk_download(ctx, hr, hv, ha);
    sgk_destroy(ctx);
    if (fileDiffuse) std::fclose(fileDiffuse);
    return 0;
}

}  // extern "C"

// C++ linkage, as the reference's definition (SurfaceGrowth.cu:2659) and its driver's call
// (SurfaceGrowth.cpp:417) have it.
void AccumProps(int icode, SimParams *hparams) {
    if (hparams) sg::accum_props(icode, *hparams);
}
