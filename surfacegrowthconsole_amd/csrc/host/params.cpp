// params.cpp - input file, parameters and units, random numbers, initial velocities.
//
// Restates the host driver of the reference (SurfaceGrowthConsole/SurfaceGrowth.cpp) with the same
// mixture of float and double arithmetic, so that region, cell grid and invWidth - and therefore
// every atom's cell index - come out bit-identical.  Compiled with -ffp-contract=off.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "host_internal.h"

namespace sg {

// Zhou et al. EAM parameters (SurfaceGrowthProto.h:31-80).  Double literals stored into `real`
// fields round decimal -> double -> float exactly as the reference's initialisers do.
const MATERIAL kMaterials[6] = {
    {"Cu", 2.556162, 1.554485, 22.150141, 7.669911, 4.090619, 0.327584, 0.468735, 0.431307,
     0.86214, {-2.176490, -0.140035, 0.285621, -1.750834}, {-2.19, 0.f, 0.702991, 0.683705},
     0.921150, -2.191675, 63.54, 8930.f},
    {"Ag", 2.891814, 1.106232, 15.539255, 7.944536, 4.237086, 0.266074, 0.386272, 0.425351,
     0.850703, {-1.729619, -0.221025, 0.541558, -0.967036}, {-1.75, 0.f, 0.983967, 0.520904},
     1.149461, -1.751274, 107.868, 10490.f},
    {"Au", 2.885034, 1.529021, 21.319637, 8.086176, 4.312627, 0.230728, 0.336695, 0.420755,
     0.841511, {-2.930281, -0.554034, 1.489437, -0.886809}, {-2.98, 0.f, 2.283863, 0.494127},
     1.286960, -2.981365, 196.9665, 19302.f},
    {"Ni", 2.488746, 2.007018, 27.984706, 8.029633, 4.282471, 0.439664, 0.632771, 0.413436,
     0.826873, {-2.693996, -0.066073, 0.170482, -2.457442}, {-2.70, 0.f, 0.282257, 0.102879},
     0.509860, -2.700493, 58.71, 8902.f},
    {"Al", 2.886166, 1.392302, 20.226537, 6.942419, 3.702623, 0.251519, 0.313394, 0.395132,
     0.790264, {-2.806783, -0.276173, 0.893409, -1.637201}, {-2.83, 0.f, 0.929508, -0.682320},
     0.779208, -2.829437, 26.98154, 2698.9},
    {"Pb", 3.499723, 0.647872, 8.906840, 8.468412, 4.516486, 0.134878, 0.203093, 0.425877,
     0.851753, {-1.419644, -0.228622, 0.630069, -0.560952}, {-1.44, 0.f, 0.921049, 0.108847},
     1.172361, -1.440494, 207.19, 11300.f},
};

static SgOptions g_options;
static bool g_options_init = false;

SgOptions &options() {
    if (!g_options_init) {
        sgDefaultOptions(&g_options);
        g_options_init = true;
    }
    return g_options;
}

}  // namespace sg

using sg::options;

extern "C" {

const char *sgVersion(void) { return "surfacegrowth-mi355x 0.1 (gfx950)"; }

int sgStructSizes(int *out, int max) {
    const int s[] = {(int)sizeof(SimParams), (int)sizeof(SgOptions), (int)sizeof(SgInput), (int)sizeof(TBuf),
                     (int)sizeof(MATERIAL)};
    int n = 0;
    for (; n < max && n < 5; ++n) out[n] = s[n];
    return n;
}

void sgDefaultOptions(SgOptions *o) {
    std::memset(o, 0, sizeof(*o));
    // SurfaceGrowth.cpp:50-51, 62-63, 81-89, 35
    o->rNebrShell = 0.4f;
    o->iNebrMax = 2 * SG_BLOCK_SIZE;
    o->blockSize = SG_BLOCK_SIZE;
    o->stepThermostat = 25;
    o->gammaBerendsen = (real)0.005;
    o->limitRdf = 100;
    o->sizeHistRdf = 200;
    o->nValDiffuse = 25000;
    o->nBuffDiffuse = 10;
    o->stepDiffuse = 100;
    o->limitDiffuseAv = 20;
    o->stepBckup = 10000;
    // extensions
    o->seed = 0;
    o->geometry = SG_GEOM_REFERENCE;
    o->gapAngstrom = 2.5f;
    o->velocityRule = SG_VEL_REFERENCE;
    o->precision = SG_PREC_FP32;
    o->rebuildEvery = 0;
    o->rangeRdfAngstrom = 0.f;
    o->stepsPerSync = 1000;
    o->useGraphs = 1;
    o->metalThermostatAll = 0;
    o->tauPs = 0.f;
}

void sgSetOptions(const SgOptions *opt) {
    sg::options() = *opt;
}

void sgGetOptions(SgOptions *opt) { *opt = sg::options(); }

void sgDefaultInput(SgInput *in) {
    // Globals of SurfaceGrowth.cpp:39-77 before the input file overrides them.
    std::memset(in, 0, sizeof(*in));
    in->iRegime = 2;
    in->iMaterial = 3;
    in->ux = in->uy = in->uz = 7;
    in->temperature = 298.f;
    in->nMolMe = 300;
    in->epsilon = (real)0.0087381;
    in->sigma = (real)2.4945;
    in->stepLimit = 50000;
    in->deltaT = (real)0.001;
    in->stepAvg = 250;
    in->stepEquil = 50000;
    in->deltaF = (real)0.0001;
    in->stepCool = 5000;
    in->deposEnergy = (real)0.03;
    in->nMolToDeposit = 40;
    in->stepDeposit = 11000;
    in->stepPdb = 500;
    in->stepRdf = 50;
    in->bBckup = 0;
    in->bStartBckup = 0;
}

// "key=value" extension rows (row >= 23); unknown keys are ignored like any later row.
static void apply_extension(const std::string &tok, SgOptions *o) {
    size_t eq = tok.find('=');
    if (eq == std::string::npos || !o) return;
    std::string k = tok.substr(0, eq), v = tok.substr(eq + 1);
    auto i = [&]() { return std::atoi(v.c_str()); };
    auto f = [&]() { return (real)std::atof(v.c_str()); };
    if (k == "seed") o->seed = i();
    else if (k == "geometry") o->geometry = v == "sphere" ? SG_GEOM_SPHERE : v == "icosahedron" ? SG_GEOM_ICOSAHEDRON : SG_GEOM_REFERENCE;
    else if (k == "gap") o->gapAngstrom = f();
    else if (k == "velocities") o->velocityRule = v == "maxwell" ? SG_VEL_MAXWELL : SG_VEL_REFERENCE;
    else if (k == "precision") o->precision = v == "mixed" ? SG_PREC_MIXED : SG_PREC_FP32;
    else if (k == "rebuild") o->rebuildEvery = i();
    else if (k == "rdf_range") o->rangeRdfAngstrom = f();
    else if (k == "rdf_bins") o->sizeHistRdf = i();
    else if (k == "rdf_limit") o->limitRdf = i();
    else if (k == "tau") o->tauPs = f();
    else if (k == "gamma") o->gammaBerendsen = f();
    else if (k == "thermostat_every") o->stepThermostat = i();
    else if (k == "steps_per_sync") o->stepsPerSync = i();
    else if (k == "graphs") o->useGraphs = i();
    else if (k == "diffuse_every") o->stepDiffuse = i();
    else if (k == "diffuse_len") o->nValDiffuse = i();
    else if (k == "diffuse_buffers") o->nBuffDiffuse = i();
    else if (k == "diffuse_avg") o->limitDiffuseAv = i();
    else if (k == "backup_every") o->stepBckup = i();
    else if (k == "metal_thermostat") o->metalThermostatAll = v == "all";
    else if (k == "skin") o->rNebrShell = f();
    else if (k == "nebr_max") o->iNebrMax = i();
}

// ReadInputFile (SurfaceGrowth.cpp:633-785).  A row is a '\n'-terminated line that is not blank
// and does not start (after spaces) with '#'.  Its value is the text from the first character of
// the line up to the first space, converted with atoi/atof (a line starting with a space therefore
// reads as 0, as in the reference).  A final line without '\n' is not read, as in the reference.
int sgReadInputFile(const char *path, SgInput *in, SgOptions *opt) {
    FILE *f = std::fopen(path, "r");
    if (!f) return 0;
    std::string line;
    int row = 0;
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c != '\n') {
            line.push_back((char)c);
            continue;
        }
        size_t k = 0;
        while (k < line.size() && line[k] == ' ') ++k;
        if (k == line.size() || line[k] == '#') {
            line.clear();
            continue;
        }
        ++row;
        size_t sp = line.find(' ');
        std::string tok = line.substr(0, sp);  // npos: whole line (the reference overruns here)
        line.clear();
        const char *t = tok.c_str();
        switch (row) {
            case 1: in->iRegime = std::atoi(t); break;
            case 2: in->iMaterial = std::atoi(t); break;
            case 3: in->ux = std::atoi(t); break;
            case 4: in->uy = std::atoi(t); break;
            case 5: in->uz = std::atoi(t); break;
            case 6: in->temperature = (real)std::atof(t); break;
            case 7: in->nMolMe = std::atoi(t); break;
            case 8: in->epsilon = (real)std::atof(t); break;
            case 9: in->sigma = (real)std::atof(t); break;
            case 10: in->stepLimit = std::atoi(t); break;
            case 11: in->deltaT = (real)std::atof(t); break;
            case 12: in->stepAvg = std::atoi(t); break;
            case 13: in->stepEquil = std::atoi(t); break;
            case 14: in->deltaF = (real)std::atof(t); break;
            case 15: in->stepCool = std::atoi(t); break;
            case 16: in->deposEnergy = (real)std::atof(t); break;
            case 17: in->nMolToDeposit = std::atoi(t); break;
            case 18: in->stepDeposit = std::atoi(t); break;
            case 19: in->stepPdb = std::atoi(t); break;
            case 20: in->stepRdf = std::atoi(t); break;
            case 21: in->bBckup = std::atoi(t); break;
            case 22: in->bStartBckup = std::atoi(t); break;
            default: apply_extension(tok, opt); break;
        }
    }
    std::fclose(f);
    return 1;
}

// Rapaport LCG (SurfaceGrowth.cpp:548-593, constants SurfaceGrowth.h:38-41).
void sgInitRand(int seed, SimParams *p) {
    if (seed != 0) {
        p->randSeedP = seed;
    } else {
        // The reference reads tm_sec of an uninitialised struct tm after mktime; any time-derived
        // value has the same role.
        p->randSeedP = (int)(std::time(nullptr) % 60);
    }
}

real sgRandR(SimParams *p) {
    unsigned int s = (unsigned int)p->randSeedP * 314159269u + 453806245u;  // IMUL, IADD
    p->randSeedP = (int)(s & 2147483647u);                                    // MASK
    return (real)(p->randSeedP * 0.4656612873e-9);                           // SCALE (double)
}

void sgVRand(sg_float3 *v, SimParams *p) {
    real s = 2., x = 0, y = 0;
    while (s > 1.) {
        x = (real)(2. * sgRandR(p) - 1.);
        y = (real)(2. * sgRandR(p) - 1.);
        s = x * x + y * y;
    }
    v->z = (real)(1. - 2. * s);
    s = (real)(2. * std::sqrt(1. - s));
    v->x = s * x;
    v->y = s * y;
}

// EamInit (SurfaceGrowth.cpp:480-529).
static void eam_init(SimParams &p, real a0) {
    const MATERIAL &m = sg::kMaterials[p.iMaterial];
    p.re = m.re / a0;
    p.a = (real)(p.re * std::sqrt(2.));
    p.fe = m.fe;
    p.rhoe = m.rhoe;
    p.alpha = m.alpha;
    p.beta = m.beta;
    p.A = m.A / p.enU;
    p.B = m.B / p.enU;
    p.kappa = m.kappa;
    p.lambda = m.lambda;
    for (int k = 0; k < 4; ++k) {
        p.Fn[k] = m.Fn[k] / p.enU;
        p.F[k] = m.F[k] / p.enU;
    }
    p.eta = m.eta;
    p.Fe = m.Fe / p.enU;
    p.massMe = m.massMe / p.massU;
    p.density = (real)(m.density * 0.1435655 * 0.001);
    p.rCutEam = (real)(1.45 * p.a);
    p.rrCutEam = p.rCutEam * p.rCutEam;
    p.rei = 1 / p.re;
    p.rhoei = 1 / p.rhoe;
}

// GrapheneInit (SurfaceGrowth.cpp:531-545).
static void graphene_init(SimParams &p, real a0, real epsilon, real sigma) {
    p.mu_r = (real)(41.881 * a0 * a0 / p.enU);
    p.theta_0 = (real)(2. * M_PI / 3.);
    p.mu_theta = (real)(2.9959 * a0 * a0 / p.enU);
    p.mu_p = (real)(18.225 * a0 * a0 / p.enU);
    p.rCutC = (real)1.2;
    p.rrCutC = p.rCutC * p.rCutC;
    p.epsLJ = epsilon / p.enU;
    p.sigmaLJ = sigma / a0;
    p.rCutLJ = (real)(2.5 * p.sigmaLJ);
    p.rrCutLJ = p.rCutLJ * p.rCutLJ;
}

static int shift_cells_shear(int n) {  // SurfaceGrowth.cpp:260-266
    if (n == 0) return 0;
    if (n < 100) return 1;
    if (n < 1000) return 2;
    if (n < 5000) return 3;
    if (n < 10000) return 5;
    if (n < 20000) return 7;
    return 9;
}

static int shift_cells_growth(int n) {  // SurfaceGrowth.cpp:241-246
    if (n == 0) return 0;
    if (n < 100) return 1;
    if (n < 5000) return 2;
    if (n < 10000) return 3;
    if (n < 20000) return 4;
    return 5;
}

// SetParams (SurfaceGrowth.cpp:189-396) without the SetParametersW call.
int sgSetParams(const SgInput *in, SimParams *pp, long long maxGridSize) {
    SimParams &p = *pp;
    const SgOptions &o = options();
    std::memset(&p, 0, sizeof(p));
    const real a0 = (real)1.42;  // g_ha0
    const real shell = o.rNebrShell;
    // units (SurfaceGrowth.cpp:192-197)
    p.temperatureU = (real)(298 / 0.409246);
    p.enU = (real)0.06275049;
    p.kB = (real)0.99999793;
    p.lengthU = a0;
    p.forceU = (real)14.1239836;
    p.massU = (real)12.0107;
    p.iRegime = in->iRegime;
    p.iMaterial = in->iMaterial < 0 || in->iMaterial > 5 ? 3 : in->iMaterial;
    std::memcpy(p.szNameMe, sg::kMaterials[p.iMaterial].szName, 2);
    int ux = in->ux < 2 ? 2 : in->ux, uy = in->uy < 2 ? 2 : in->uy, uz = in->uz;
    if (p.iRegime == 0) {
        p.nMolMe = 4 * ux * uy * uz;
        p.nMol = p.nMolMe;
    } else {
        p.nMolMe = in->nMolMe;
        p.nMol = p.nMolMe + 32 * ux * uy;
        uz = 0;
    }
    eam_init(p, a0);
    graphene_init(p, a0, in->epsilon, in->sigma);
    float rx, ry, rz;
    int cellShiftZ = 4;  // g_hcellShiftZ default (SurfaceGrowth.cpp:64)
    const real w = p.rCutEam + shell;
    if (p.iRegime == 0) {
        double s = 1. / std::pow((p.density * 0.25 / p.massMe), 1. / 3.);
        rx = (float)(s * ux);
        ry = (float)(s * uy);
        rz = (float)(s * uz);
    } else {
        rx = (float)(ux * 8 * std::cos(M_PI / 6.));
        ry = (float)(uy * 6);
        rz = 0.f;
        if (p.iRegime == 1) {
            cellShiftZ = shift_cells_growth(p.nMolMe);
            if (p.nMolMe != 0)
                rz = (float)(1.8 * p.a * std::pow((0.25 * p.nMolMe), 0.33333333333) + p.a + cellShiftZ * w);
            else
                rz = 3 * w;
        }
        if (p.iRegime == 2) {
            cellShiftZ = shift_cells_shear(p.nMolMe);
            if (p.nMolMe != 0)
                rz = (float)(p.a * std::pow((0.25 * p.nMolMe), 0.33333333333) + 3.6 * p.a + cellShiftZ * w);
            else
                rz = 3 * w;
            // Extension: a sphere or icosahedron is taller than the cube the reference's height
            // rule assumes.  Raise the box so the cluster top stays one cell below the boundary.
            if (o.geometry != SG_GEOM_REFERENCE && p.nMolMe != 0) {
                double zc0 = -0.5 * rz + cellShiftZ * w;  // z_0 estimate with cell width ~ w
                double diam;
                if (o.geometry == SG_GEOM_SPHERE)
                    diam = 2.0 * std::cbrt(3.0 * p.nMolMe / (16.0 * M_PI)) * p.a + p.a;
                else
                    diam = 2.0 * (sg::icosahedron_shells(p.nMolMe) + 1) * p.re;
                double top = zc0 + o.gapAngstrom / a0 + diam + 2.0 * w;
                if (top > 0.5 * rz) rz = (float)(rz + 2.0 * (top - 0.5 * rz));
            }
        }
    }
    const double inv = 1.0 / w;
    p.cells.x = (int)(inv * rx);
    p.cells.y = (int)(inv * ry);
    p.cells.z = (int)(inv * rz);
    p.initUcell.x = (unsigned)ux;
    p.initUcell.y = (unsigned)uy;
    p.initUcell.z = (unsigned)uz;
    p.region.x = rx;
    p.region.y = ry;
    p.region.z = rz;
    p.invWidth.x = p.cells.x / p.region.x;
    p.invWidth.y = p.cells.y / p.region.y;
    p.invWidth.z = p.cells.z / p.region.z;
    p.cellShiftZ = cellShiftZ;
    if (p.nMolMe != 0)
        p.z_0 = -0.5f * p.region.z + p.cellShiftZ / p.invWidth.z;
    else
        p.z_0 = 0;
    real T = in->temperature <= 0 ? (real)298 : in->temperature;
    p.temperature = T / p.temperatureU;
    p.stepLimit = in->stepLimit;
    p.stepAvg = in->stepAvg <= 0 ? 1 : in->stepAvg;
    p.stepEquil = in->stepEquil;
    p.stepPdb = in->stepPdb;
    p.moreCycles = 1;
    p.deltaT = in->deltaT;
    p.rNebrShell = shell;
    p.blockSize = (unsigned)o.blockSize;
    p.maxMol = p.cells.x * p.cells.y * p.cells.z * (int)p.blockSize;
    p.gridSize = (unsigned)(p.cells.x * p.cells.y * p.cells.z);
    if (maxGridSize > 0 && (long long)p.gridSize > maxGridSize) return 0;
    p.rrNebr = (p.rCutEam + p.rNebrShell) * (p.rCutEam + p.rNebrShell);
    p.iNebrMax = o.iNebrMax;
    p.bPdb = 1;
    p.bResult = 1;
    p.stepThermostat = o.stepThermostat;
    p.gammaBerendsen = o.gammaBerendsen;
    if (o.tauPs > 0) {  // extension: Berendsen tau -> gamma = stepThermostat*dt/tau (dt in ps)
        double dtPs = p.deltaT * 0.2;  // time unit 0.2 ps (SurfaceGrowth.h:239-241)
        p.gammaBerendsen = (real)(p.stepThermostat * dtPs / o.tauPs);
    }
    p.stepCool = p.iRegime == 1 ? p.stepLimit : in->stepCool;
    p.deltaF = (real)(in->deltaF * p.forceU * 0.001);
    p.shear = 0.f;
    p.velMagDepos = std::sqrt(2 * in->deposEnergy / (p.enU * p.massMe));
    p.stepDeposit = in->stepDeposit;
    p.nMolToDeposit = in->nMolToDeposit;
    p.nMolDeposited = p.iRegime == 1 ? 0 : p.nMolMe;
    if (p.iRegime == 0)
        p.velMag = (real)std::sqrt(3 * (1. - 1. / p.nMol) * p.kB * p.temperature / p.massMe);
    else
        p.velMag = (real)std::sqrt(3 * (1. - 1. / (p.nMol)) * p.kB * p.temperature * p.nMol /
                                   (p.nMol - p.nMolMe));
    // RDF (SurfaceGrowth.cpp:365-372); range extension in angstrom
    p.rangeRdf = o.rangeRdfAngstrom > 0 ? o.rangeRdfAngstrom / a0 : (real)(1.5 * p.rCutEam);
    p.limitRdf = o.limitRdf;
    p.sizeHistRdf = o.sizeHistRdf;
    p.stepRdf = in->stepRdf;
    p.countRdf = 0;
    p.intervalRdf = p.sizeHistRdf / p.rangeRdf;
    p.bRdf = 1;
    std::strcpy(p.szRdfPath, "rdf");
    std::strcpy(p.szBckup0, "bckup0.sugr");
    std::strcpy(p.szBckup1, "bckup1.sugr");
    p.bBckup = in->bBckup;
    p.bStartBckup = in->bStartBckup;
    p.stepBckup = o.stepBckup;
    p.totalTime = 0.f;
    p.nValDiffuse = o.nValDiffuse;
    p.nBuffDiffuse = o.nBuffDiffuse;
    p.stepDiffuse = o.stepDiffuse;
    p.limitDiffuseAv = o.limitDiffuseAv;
    p.countDiffuseAv = 0;
    sgInitRand(o.seed, &p);
    return 1;
}

// Gaussian deviate from the LCG (Box-Muller), for the Maxwell-Boltzmann extension.
static double gauss(SimParams *p) {
    double u1, u2;
    do {
        u1 = sgRandR(p);
    } while (u1 <= 0.0);
    u2 = sgRandR(p);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// InitVels (SurfaceGrowth.cpp:596-621); Maxwell-Boltzmann when the extension selects it.
void sgInitVels(SimParams *p, sg_float3 *hv) {
    const SgOptions &o = options();
    if (o.velocityRule == SG_VEL_MAXWELL) {
        double P[3] = {0, 0, 0}, M = 0;
        for (int n = 0; n < p->nMol; ++n) {
            const bool metal = n < p->nMolMe;
            const double m = metal ? p->massMe : 1.0;
            const double s = std::sqrt((double)p->kB * p->temperature / m);
            double v[3] = {s * gauss(p), s * gauss(p), s * gauss(p)};
            hv[n].x = (real)v[0];
            hv[n].y = (real)v[1];
            hv[n].z = (real)v[2];
            for (int d = 0; d < 3; ++d) P[d] += m * v[d];
            M += m;
        }
        for (int n = 0; n < p->nMol; ++n) {  // zero total momentum
            hv[n].x = (real)(hv[n].x - P[0] / M);
            hv[n].y = (real)(hv[n].y - P[1] / M);
            hv[n].z = (real)(hv[n].z - P[2] / M);
        }
        return;
    }
    sg_float3 vSum = {0, 0, 0};
    for (int n = 0; n < p->nMol; ++n) {
        if (p->iRegime != 0 && n < p->nMolMe) {
            hv[n].x = hv[n].y = hv[n].z = 0.;
        } else {
            sgVRand(&hv[n], p);
            hv[n].x *= p->velMag;
            hv[n].y *= p->velMag;
            hv[n].z *= p->velMag;
            vSum.x = vSum.x + hv[n].x;
            vSum.y = vSum.y + hv[n].y;
            vSum.z = vSum.z + hv[n].z;
        }
    }
    for (int n = 0; n < p->nMol; ++n) {
        if (p->iRegime == 0 || n >= p->nMolMe) {
            hv[n].x = (real)(hv[n].x + (-1. / (p->nMol)) * vSum.x);
            hv[n].y = (real)(hv[n].y + (-1. / (p->nMol)) * vSum.y);
            hv[n].z = (real)(hv[n].z + (-1. / (p->nMol)) * vSum.z);
        }
    }
}

// Result file name and header (SurfaceGrowth.cpp:421-469).
FILE *sgOpenResultFile(const SimParams *p, real deposEnergy, char *nameOut) {
    char buf[MAX_PATH], name[2 * MAX_PATH] = "sg";
    buf[0] = 0;
    const double T = (double)(p->temperature * p->temperatureU);
    if (p->iRegime == 0)
        std::snprintf(buf, sizeof(buf), "_blk_x%i_y%i_z%i_Me%i_Av%i_Pdb%i_T%3.0f", p->initUcell.x,
                      p->initUcell.y, p->initUcell.z, p->nMolMe, p->stepAvg, p->stepPdb, T);
    if (p->iRegime == 1)
        std::snprintf(buf, sizeof(buf), "_x%i_y%i_Me%i_Eq%i_Dep%i_TD%i_Pdb%i_T%3.0f", p->initUcell.x,
                      p->initUcell.y, p->nMolMe, p->stepEquil, p->stepDeposit, p->nMolToDeposit,
                      p->stepPdb, T);
    if (p->iRegime == 2)
        std::snprintf(buf, sizeof(buf), "_sh_x%i_y%i_Me%i_Eq%i_C%i_Av%i_Pdb%i_T%3.0f", p->initUcell.x,
                      p->initUcell.y, p->nMolMe, p->stepEquil, p->stepCool, p->stepAvg, p->stepPdb, T);
    std::strncat(buf, p->szNameMe, sizeof(buf) - std::strlen(buf) - 1);
    std::strncat(buf, ".txt", sizeof(buf) - std::strlen(buf) - 1);
    std::strncat(name, buf, sizeof(name) - std::strlen(name) - 1);
    if (nameOut) std::strcpy(nameOut, name);
    FILE *f = std::fopen(name, "w");
    if (!f) return nullptr;
    std::fprintf(f,
                 "stepCnt\t impulse\t totEn(eV)\t totEn.rms(eV)\t potEn(eV)\t potEn.rms(eV)\t Tempr(K)\t "
                 "T.rms(K)\t oneStep(ms)\t Veloc_CM\t CM(angstr)\t friction(nN)\t sizex(angstr)\t "
                 "sizey(angstr)\t sizez(angstr)\t shearForce(nN)\t");
    if (p->bResult != 0) std::fprintf(f, "time step = %f\t", p->deltaT);
    if (p->bResult != 0 && p->iRegime == 1) std::fprintf(f, "deposit energy = %f eV\t", deposEnergy);
    if (p->bResult != 0 && p->iRegime == 2)
        std::fprintf(f, "increment of shear = %f pN\t", p->deltaF * 1000 / p->forceU);
    if (p->bResult != 0 && p->iRegime != 0) {
        std::fprintf(f, "epsilonLJ = %f eV\t", p->epsLJ * p->enU);
        std::fprintf(f, "sigmaLJ = %f angstrom", p->sigmaLJ * p->lengthU);
    }
    if (p->bResult != 0) std::fprintf(f, "\n");
    return f;
}

// Backup with the reference layout: raw SimParams, r float4[N], v float3[N], a float3[N]
// (SurfaceGrowth.cu:2475-2478).
int sgWriteBackup(const char *path, const SimParams *p, const sg_float4 *hr, const sg_float3 *hv,
                  const sg_float3 *ha) {
    FILE *f = std::fopen(path, "w+b");
    if (!f) return 0;
    size_t ok = std::fwrite(p, sizeof(SimParams), 1, f);
    ok += std::fwrite(hr, sizeof(sg_float4), (size_t)p->nMol, f);
    ok += std::fwrite(hv, sizeof(sg_float3), (size_t)p->nMol, f);
    ok += std::fwrite(ha, sizeof(sg_float3), (size_t)p->nMol, f);
    std::fclose(f);
    return ok == 1 + 3 * (size_t)p->nMol;
}

// Restart reader for the same layout (the reference parses bStartBckup but never reads a backup).
int sgReadBackup(const char *path, SimParams *p, sg_float4 *hr, sg_float3 *hv, sg_float3 *ha,
                 int capacity) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return 0;
    SimParams tmp;
    if (std::fread(&tmp, sizeof(SimParams), 1, f) != 1 || tmp.nMol < 0 || tmp.nMol > capacity) {
        std::fclose(f);
        return 0;
    }
    size_t n = (size_t)tmp.nMol, ok = 0;
    ok += std::fread(hr, sizeof(sg_float4), n, f);
    ok += std::fread(hv, sizeof(sg_float3), n, f);
    ok += std::fread(ha, sizeof(sg_float3), n, f);
    std::fclose(f);
    if (ok != 3 * n) return 0;
    *p = tmp;
    return 1;
}

}  // extern "C"
