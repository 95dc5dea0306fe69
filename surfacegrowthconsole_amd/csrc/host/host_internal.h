// host_internal.h - declarations shared by the host-side C++ of libsurfacegrowth.so.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "surfacegrowth.h"

namespace sg {

// EAM material table of Zhou et al. (SurfaceGrowthProto.h:31-80), in the reference's units
// (angstrom, eV, amu, kg/m^3) and order Cu, Ag, Au, Ni, Al, Pb.
extern const MATERIAL kMaterials[6];

// Process-wide options (the reference keeps them in globals, SurfaceGrowth.cpp:49-90).
SgOptions &options();

// Coordinate generators; positions in reduced units (length 1.42 A), w = 0.
// InitFccCoordsK (SurfaceGrowth.cu:223-255)
void gen_fcc_bulk(const SimParams &p, sg_float4 *r);
// InitGrapheneCoordsK (SurfaceGrowth.cu:305-394): writes r[nMolMe .. nMol)
void gen_graphene(const SimParams &p, sg_float4 *r);
// InitSlabCoordsK (SurfaceGrowth.cu:258-303) with the launch shape of SurfaceGrowth.cu:2583-2594
void gen_slab(SimParams &p, sg_float4 *r);
// Host random metal placement of regime 1 (SurfaceGrowth.cu:2561-2569)
void gen_deposit_random(SimParams &p, sg_float4 *r);
// Extensions: fcc sphere (nMolMe sites nearest a lattice site) and Mackay icosahedron, placed with
// the lowest atom `gap` (reduced) above z_0 and centred at x = y = 0.
int gen_sphere(const SimParams &p, double gap, sg_float4 *r);
int gen_icosahedron(const SimParams &p, double gap, sg_float4 *r);
// Mackay icosahedron magic number test: returns the shell count k if n = 1 + sum(10j^2+2), else -1.
int icosahedron_shells(int n);

// Output writers with the reference formats.
void accum_props(int icode, SimParams &p);                              // SurfaceGrowth.cu:2659-2686
void print_summary(FILE *fp, const SimParams &p);                       // SurfaceGrowth.cu:2689-2731
int create_pdb_file(const char *prefix, const SimParams &p, const sg_float4 *r);  // SurfaceGrowth.cu:2734-2810
std::string pdb_file_name(const char *prefix, const SimParams &p);
int print_rdf(const SimParams &p, const unsigned int *hist);            // SurfaceGrowth.cu:2812-2844
std::string rdf_file_name(const SimParams &p);

// Centre-of-mass diffusion with Rapaport's staggered buffers (SurfaceGrowth.cu:2865-2952), FP32 as
// in the reference.
class Diffusion {
public:
    Diffusion(const SimParams &p);
    // EvalDiffusion + AccumDiffusion; prints a table to `file` when an average completes.
    void eval(SimParams &p, sg_float3 com, FILE *file);
    int tables_printed() const { return printed_; }

private:
    struct Buf {
        sg_float3 orgR{}, rTrue{};
        std::vector<real> rr;
        int count = 0;
    };
    std::vector<Buf> buf_;
    std::vector<real> av_;
    int printed_ = 0;
};

}  // namespace sg
