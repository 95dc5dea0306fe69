// main.cpp - console driver with the reference's program flow (SurfaceGrowthConsole/SurfaceGrowth.cpp:104-186):
// read sugr_in.txt from the working directory, set parameters, build the initial state, run, write
// sg_*.txt / sugr_*.pdb / rdf_*.txt / Diffuse.txt.  The reference waits for a key press at the end
// (_getch); this driver exits instead.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "host_internal.h"
#include "sgk.h"

int main(int argc, char **argv) {
    const char *input = argc > 1 ? argv[1] : "sugr_in.txt";
    if (sgk_device_count() == 0) {
        std::printf("There are no HIP capable devices!\n");
        return 1;
    }
    SgInput in;
    sgDefaultInput(&in);
    SgOptions opt;
    sgDefaultOptions(&opt);
    if (sgReadInputFile(input, &in, &opt) == 0) {
        std::printf("Error with input file! Check it!\n");
        return 1;
    }
    sgSetOptions(&opt);
    CudaInitW(argc, argv);
    SimParams p;
    if (!sgSetParams(&in, &p, 0)) {
        std::printf("The system is too large (grid = %i blocks)!\n", p.gridSize);
        std::printf("The application cannot be launched! Exit!\n");
        return 1;
    }
    SetParametersW(&p);
    // SetupJob (SurfaceGrowth.cpp:399-477)
    std::vector<sg_float4> hr(p.nMol);
    std::vector<sg_float3> hv(p.nMol), ha(p.nMol, sg_float3{0.f, 0.f, 0.f});
    p.stepCount = 0;
    const char *err = InitCoordsW(nullptr, hr.data(), &p);
    if (err) {
        std::printf("Exception: %s\n", err);
        std::printf("Problems with initial coordinates! We exit!\n");
        return 1;
    }
    sgInitVels(&p, hv.data());
    sg::accum_props(0, p);
    p.nebrNow = 1;
    char resultName[2 * MAX_PATH];
    FILE *fResult = sgOpenResultFile(&p, in.deposEnergy, resultName);
    if (p.iRegime == 2) std::snprintf(p.szDiffusePath, MAX_PATH, "%s", "Diffuse.txt");
    std::printf("Performing computations. Wait...\n");
    char szPdb[MAX_PATH] = "sugr";
    char *e = DoComputationsW(hr.data(), hv.data(), ha.data(), &p, fResult, szPdb);
    if (e != 0) {
        std::printf("%s\n", e);
        std::printf("Problems with DoComputationsW! We exit!\n");
        if (fResult) std::fclose(fResult);
        return 1;
    }
    if (fResult) std::fclose(fResult);
    std::printf("Done!\n");
    return 0;
}
