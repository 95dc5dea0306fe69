// output.cpp - property accumulation and the output files of the reference.
//
// Formats and float arithmetic follow SurfaceGrowthConsole/SurfaceGrowth.cu:2659-2952 so that the
// files sg_*.txt, sugr_*.pdb, rdf_stepCount_*_<Me>.txt and Diffuse.txt have the reference's layout.
#include <cmath>
#include <cstring>
#include <string>

#include "host_internal.h"

namespace sg {

static inline void prop_zero(Prop &v) { v.sum = v.sum2 = 0.; }
static inline void prop_accum(Prop &v) {
    v.sum += v.val;
    v.sum2 += v.val * v.val;
}
static inline void prop_avg(Prop &v, int n) {  // PropAvg (ComputeDefs.h:56-57)
    v.sum /= n;
    float t = v.sum2 / n - v.sum * v.sum;
    v.sum2 = (real)std::sqrt(t > 0. ? (double)t : 0.);
}

// AccumProps (SurfaceGrowth.cu:2659-2686).
void accum_props(int icode, SimParams &p) {
    Prop *props[] = {&p.totEnergy, &p.potEnergy, &p.kinEnergy, &p.oneStep, &p.centerOfMass, &p.cmVel, &p.frictForce};
    for (Prop *v : props) {
        if (icode == 0)
            prop_zero(*v);
        else if (icode == 1)
            prop_accum(*v);
        else if (icode == 2)
            prop_avg(*v, p.stepAvg);
    }
}

// PrintSummary (SurfaceGrowth.cu:2689-2731).
void print_summary(FILE *fp, const SimParams &p) {
    real totEn = p.totEnergy.sum * p.enU;
    real totEnRms = p.totEnergy.sum2 * p.enU;
    real potEn = p.potEnergy.sum * p.enU;
    real potEnRms = p.potEnergy.sum2 * p.enU;
    real T = 2.f * p.kinEnergy.sum * p.temperatureU / (3 * p.kB);
    real TRms = 2.f * p.kinEnergy.sum2 * p.temperatureU / (3 * p.kB);
    real impulse = (p.vSum.x + p.vSum.y + p.vSum.z) / p.nMol;
    std::fprintf(fp, "%5d\t %7.7f\t %7.7f\t %7.7f\t %7.7f\t %7.7f\t %7.7f\t %7.7f\t %7.7f\t", p.stepCount,
                 impulse, totEn, totEnRms, potEn, potEnRms, T, TRms, p.oneStep.sum);
    real centerOfMass = p.centerOfMass.sum * p.lengthU;
    real frictForce = p.frictForce.sum / p.forceU;
    real sx = p.particleSize.x * p.lengthU, sy = p.particleSize.y * p.lengthU, sz = p.particleSize.z * p.lengthU;
    real totalShear = p.totalShear / p.forceU;
    std::fprintf(fp, "%7.7f\t %7.7f\t %7.6f\t %7.7f\t %7.7f\t %7.7f\t %7.7f\n", p.cmVel.sum, centerOfMass,
                 frictForce, sx, sy, sz, totalShear);
    std::fflush(fp);
}

// File name of CreatePdbFile (SurfaceGrowth.cu:2741-2754): no suffix at or beyond 10000.
std::string pdb_file_name(const char *prefix, const SimParams &p) {
    char buf[MAX_PATH];
    buf[0] = 0;
    int i = p.stepCount / p.stepPdb;
    if (i < 100)
        std::snprintf(buf, sizeof(buf), "_%i", i);
    else if (i < 1000)
        std::snprintf(buf, sizeof(buf), "_%-3i", i);
    else if (i < 10000)
        std::snprintf(buf, sizeof(buf), "_%-4i", i);
    return std::string(prefix) + buf + ".pdb";
}

// CreatePdbFile (SurfaceGrowth.cu:2734-2810).
int create_pdb_file(const char *prefix, const SimParams &p, const sg_float4 *r) {
    std::string name = pdb_file_name(prefix, p);
    FILE *pdb = std::fopen(name.c_str(), "w");
    if (!pdb) return 0;
    const float m = 0.f;
    const int n = 0;
    for (int i = 0; i < p.nMol; ++i) {
        const bool carbon = i >= p.nMolMe;
        std::fprintf(pdb, "ATOM  ");
        std::fprintf(pdb, "%-7i", i);
        if (carbon)
            std::fprintf(pdb, "C  ");
        else
            std::fprintf(pdb, "%2s ", p.szNameMe);
        std::fprintf(pdb, " ");
        if (carbon)
            std::fprintf(pdb, " C   ");
        else
            std::fprintf(pdb, " %2s  ", p.szNameMe);
        std::fprintf(pdb, " ");
        std::fprintf(pdb, "    ");
        std::fprintf(pdb, "%-4i", n);
        std::fprintf(pdb, "%-8.2lf", 1.42 * r[i].x);
        std::fprintf(pdb, "%-8.2lf", 1.42 * r[i].y);
        if (p.iRegime == 1 && i < p.nMolMe && i >= p.nMolDeposited)
            std::fprintf(pdb, "%-8.2lf", 1.42 * 2 * p.region.z);
        else
            std::fprintf(pdb, "%-8.2lf", 1.42 * r[i].z);
        std::fprintf(pdb, " %4.1lf", (double)m);
        std::fprintf(pdb, "  1.00");
        std::fprintf(pdb, "          ");
        if (carbon)
            std::fprintf(pdb, " C");
        else
            std::fprintf(pdb, "%2s", p.szNameMe);
        std::fprintf(pdb, "  \n");
    }
    std::fclose(pdb);
    return 1;
}

std::string rdf_file_name(const SimParams &p) {
    char buf[MAX_PATH];
    std::snprintf(buf, sizeof(buf), "_stepCount_%i_", p.stepCount);
    return std::string(p.szRdfPath) + buf + p.szNameMe + ".txt";
}

// PrintRdf (SurfaceGrowth.cu:2812-2844), called when countRdf == limitRdf.
int print_rdf(const SimParams &p, const unsigned int *hist) {
    std::string name = rdf_file_name(p);
    real normFac = (real)((p.particleSize.x * p.particleSize.y * p.particleSize.z) *
                          (p.intervalRdf * p.intervalRdf * p.intervalRdf) / (2.f * M_PI * p.countRdf));
    normFac = normFac / p.nMolMe;
    normFac = normFac / p.nMolMe;
    FILE *f = std::fopen(name.c_str(), "w+");
    if (!f) return 0;
    for (int n = 0; n < p.sizeHistRdf; ++n) {
        real rb = (n + 0.5f) * p.rangeRdf * p.lengthU / p.sizeHistRdf;
        real g = (real)hist[n] * normFac / ((n - 0.5f) * (n - 0.5f));
        std::fprintf(f, "%8.4f\t %8.4f\n", rb, g);
    }
    std::fclose(f);
    return 1;
}

// InitDiffusion (SurfaceGrowth.cu:2865-2872).
Diffusion::Diffusion(const SimParams &p) : buf_(p.nBuffDiffuse), av_(p.nValDiffuse, 0.f) {
    for (int nb = 0; nb < p.nBuffDiffuse; ++nb) {
        buf_[nb].rr.assign(p.nValDiffuse, 0.f);
        buf_[nb].count = -nb * p.nValDiffuse / p.nBuffDiffuse;
    }
}

static inline int nint(float x) {  // Nint (ComputeDefs.h:64-65), double math via its 0.5 literal
    return (x < 0.) ? -(int)(0.5 - x) : (int)(0.5 + x);
}

// EvalDiffusion + AccumDiffusion + PrintDiffusion (SurfaceGrowth.cu:2884-2952).
void Diffusion::eval(SimParams &p, sg_float3 com, FILE *file) {
    for (auto &b : buf_) {
        if (b.count == 0) {
            b.orgR = com;
            b.rTrue = com;
        }
        if (b.count >= 0) {
            const int ni = b.count;
            b.rr[ni] = 0.;
            sg_float3 dr = {b.rTrue.x - com.x, b.rTrue.y - com.y, b.rTrue.z - com.z};
            dr.x = dr.x / p.region.x;
            dr.y = dr.y / p.region.y;
            dr.z = dr.z / p.region.z;
            dr.x = (float)nint(dr.x);
            dr.y = (float)nint(dr.y);
            dr.z = (float)nint(dr.z);
            dr.x = dr.x * p.region.x;
            dr.y = dr.y * p.region.y;
            dr.z = dr.z * p.region.z;
            b.rTrue = {com.x + dr.x, com.y + dr.y, com.z + dr.z};
            dr = {b.rTrue.x - b.orgR.x, b.rTrue.y - b.orgR.y, b.rTrue.z - b.orgR.z};
            b.rr[ni] += dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
        }
        ++b.count;
    }
    for (auto &b : buf_) {
        if (b.count != p.nValDiffuse) continue;
        for (int j = 0; j < p.nValDiffuse; ++j) av_[j] += b.rr[j];
        b.count = 0;
        ++p.countDiffuseAv;
        if (p.countDiffuseAv == p.limitDiffuseAv) {
            real fac = (real)(1. / (3 * 2 * p.stepDiffuse * p.deltaT * p.limitDiffuseAv));
            for (int j = 1; j < p.nValDiffuse; ++j) av_[j] *= fac / j;
            if (file) {
                std::fprintf(file, "diffusion\n");
                for (int j = 0; j < p.nValDiffuse; ++j) {
                    real tVal = j * p.stepDiffuse * p.deltaT;
                    std::fprintf(file, "%8.4f %8.4f\n", tVal, av_[j]);
                }
            }
            ++printed_;
            p.countDiffuseAv = 0;  // ZeroDiffusion (SurfaceGrowth.cu:2875-2881)
            for (auto &v : av_) v = 0.;
        }
    }
}

}  // namespace sg

struct sgDiffusion {
    sg::Diffusion d;
    explicit sgDiffusion(const SimParams &p) : d(p) {}
};

extern "C" {

sgDiffusion *sgDiffusionCreate(const SimParams *p) {
    if (!p || p->nValDiffuse <= 0 || p->nBuffDiffuse <= 0) return nullptr;
    return new sgDiffusion(*p);
}

int sgDiffusionEval(sgDiffusion *d, SimParams *p, const sg_float3 *com, FILE *file) {
    if (!d || !p || !com) return -1;
    d->d.eval(*p, *com, file);
    return d->d.tables_printed();
}

void sgDiffusionDestroy(sgDiffusion *d) { delete d; }

}  // extern "C"
