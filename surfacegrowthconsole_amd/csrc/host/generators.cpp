// generators.cpp - initial coordinates.
//
// The reference computes its initial coordinates with three CUDA kernels and one host loop
// (SurfaceGrowthConsole/SurfaceGrowth.cu:223-394, 2532-2609).  They run once, so here they are host
// code; each restates the kernel's float arithmetic per thread so the coordinates (and the exact
// float compares of the frozen graphene perimeter, SurfaceGrowth.cu:416-424) are bit-identical.
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "host_internal.h"

namespace sg {

// InitFccCoordsK (SurfaceGrowth.cu:223-255) over grid (ux, uy) x block (1, 1, uz).
void gen_fcc_bulk(const SimParams &p, sg_float4 *r) {
    const unsigned gx = p.initUcell.x, gy = p.initUcell.y, gz = p.initUcell.z;
    const float gapx = p.region.x / p.initUcell.x, gapy = p.region.y / p.initUcell.y,
                gapz = p.region.z / p.initUcell.z;
    for (unsigned nz = 0; nz < gz; ++nz)
        for (unsigned ny = 0; ny < gy; ++ny)
            for (unsigned nx = 0; nx < gx; ++nx) {
                unsigned n = 4 * (nx + ny * gx + nz * gx * gy);
                if ((long long)n > (long long)p.nMol - 4) continue;
                float cx = (nx + 0.25f) * gapx + (-0.5f) * p.region.x;
                float cy = (ny + 0.25f) * gapy + (-0.5f) * p.region.y;
                float cz = (nz + 0.25f) * gapz + (-0.5f) * p.region.z;
                for (int j = 0; j < 4; ++j, ++n) {
                    sg_float4 q = {cx, cy, cz, 0.f};
                    if (j != 3) {
                        if (j != 0) q.x += 0.5f * gapx;
                        if (j != 1) q.y += 0.5f * gapy;
                        if (j != 2) q.z += 0.5f * gapz;
                    }
                    r[n] = q;
                }
            }
}

// InitGrapheneCoordsK (SurfaceGrowth.cu:305-394): 8-atom motifs, 32 atoms (4 motifs) per block,
// blocks laid out in rows of initUcell.x.  Writes r[nMolMe + n].
void gen_graphene(const SimParams &p, sg_float4 *r) {
    const float cos30 = 0.8660254038f;
    const int nC = p.nMol - p.nMolMe;
    const int numBlocks = nC / 32;
    const unsigned ucx = p.initUcell.x;
    for (int b = 0; b < numBlocks; ++b) {
        for (unsigned t = 0; t < 32; ++t) {
            const int n = b * 32 + (int)t;
            if (n >= nC) continue;
            int cy = (int)((unsigned)b / ucx);
            int cx = (int)((unsigned)b - (unsigned)cy * ucx);
            float dx = cx * 8.f * cos30;
            float dy = cy * 6.f;
            cy = (int)(t / 8.f);
            cx = (int)(t - cy * 8.f);
            float x = 0.f, y = 0.f;
            switch (cx) {
                case 0: x = cos30; y = 0.f; break;
                case 1: x = 0.f; y = 0.5f; break;
                case 2: x = 0.f; y = 1.5f; break;
                case 3: x = cos30; y = 2.f; break;
                case 4: x = 2.f * cos30; y = 1.5f; break;
                case 5: x = 2.f * cos30; y = 0.5f; break;
                case 6: x = 3.f * cos30; y = 2.f; break;
                case 7: x = 3.f * cos30; y = 0.f; break;
            }
            if (cy == 1) {
                dx += 4.f * cos30;
            } else if (cy == 2) {
                dy += 3.f;
            } else if (cy == 3) {
                dx += 4.f * cos30;
                dy += 3.f;
            }
            dx = dx + 0.5f * (cos30 - p.region.x);
            dy = dy + 0.5f * (1.f - p.region.y);
            r[n + p.nMolMe] = {x + dx, y + dy, p.z_0, 0.f};
        }
    }
}

// InitSlabCoordsK (SurfaceGrowth.cu:258-303) with the launch shape and bookkeeping of
// SurfaceGrowth.cu:2581-2597.
void gen_slab(SimParams &p, sg_float4 *r) {
    if (p.nMolMe == 0) return;
    const int ucx = (int)((p.region.x - 2 * p.a) / p.a);
    const int ucy = (int)((p.region.y - 2 * p.a) / p.a);
    const int layers = (int)std::ceil(p.nMolMe * 0.25 / (ucx * ucy));
    p.initUcellMeX = ucx;
    p.initUcellMeY = ucy;
    const unsigned gx = (unsigned)ucx, gy = (unsigned)ucy;
    const float rgx = gx * p.a, rgy = gy * p.a, rgz = p.region.z;
    const float shiftY = layers == 1 ? 0.5f * (rgy - std::ceil(0.25f * p.nMolMe / gx) * p.a) : 0.f;
    const float gap = p.a;
    for (unsigned nz = 0; nz < (unsigned)layers; ++nz)
        for (unsigned ny = 0; ny < gy; ++ny)
            for (unsigned nx = 0; nx < gx; ++nx) {
                unsigned n = 4 * (nx + ny * gx + nz * gx * gy);
                float cx = (nx + 0.25f) * gap + (-0.5f) * rgx;
                float cy = (ny + 0.25f) * gap + (-0.5f) * rgy;
                float cz = (nz + 0.25f) * gap + (-0.5f) * rgz;
                for (int j = 0; j < 4; ++j) {
                    if ((int)n < p.nMolMe) {
                        sg_float4 q;
                        q.x = cx;
                        q.y = cy + shiftY;
                        q.z = cz + 0.5f * p.a + p.cellShiftZ / p.invWidth.z;
                        q.w = 0.f;
                        if (j != 3) {
                            if (j != 0) q.x += 0.5f * gap;
                            if (j != 1) q.y += 0.5f * gap;
                            if (j != 2) q.z += 0.5f * gap;
                        }
                        r[n] = q;
                        n += 1;
                    }
                }
            }
}

// Regime-1 metal atoms parked above the box (SurfaceGrowth.cu:2561-2569, VRandRfloat4 2847-2861).
void gen_deposit_random(SimParams &p, sg_float4 *r) {
    for (int i = 0; i < p.nMolMe; ++i) {
        float s = 2.f, x = 0.f, y = 0.f;
        while (s > 1.f) {
            x = 2.f * sgRandR(&p) - 1.f;
            y = 2.f * sgRandR(&p) - 1.f;
            s = x * x + y * y;
        }
        float z = 1.f - 2.f * s;
        (void)z;
        s = 2.f * std::sqrt(1.f - s);
        float vx = s * x, vy = s * y;
        r[i].x = 0.5f * p.region.x * vx;
        r[i].y = 0.5f * p.region.y * vy;
        r[i].z = 1.5f * p.region.z + 0.5f * (i + 1) * p.region.z;
        r[i].w = 0.f;
    }
}

// Place a cluster given in double coordinates: lowest atom `gap` above z_0, centroid at x=y=0,
// atoms ordered by layer (z, then y, then x) so neighbours are close in memory.
static void place_cluster(const SimParams &p, double gap, std::vector<std::array<double, 3>> &pts,
                          sg_float4 *r) {
    double cx = 0, cy = 0, zmin = 1e300;
    for (auto &q : pts) {
        cx += q[0];
        cy += q[1];
        zmin = std::min(zmin, q[2]);
    }
    cx /= (double)pts.size();
    cy /= (double)pts.size();
    std::stable_sort(pts.begin(), pts.end(), [](const std::array<double, 3> &a, const std::array<double, 3> &b) {
        const double eps = 1e-9;
        if (std::fabs(a[2] - b[2]) > eps) return a[2] < b[2];
        if (std::fabs(a[1] - b[1]) > eps) return a[1] < b[1];
        return a[0] < b[0];
    });
    const double dz = (double)p.z_0 + gap - zmin;
    for (size_t n = 0; n < pts.size(); ++n)
        r[n] = {(float)(pts[n][0] - cx), (float)(pts[n][1] - cy), (float)(pts[n][2] + dz), 0.f};
}

// Sphere cut from the fcc lattice of the metal (lattice constant p.a): the nMolMe lattice sites
// nearest to a lattice site, ties broken by lattice coordinates.  201 is a closed shell.
int gen_sphere(const SimParams &p, double gap, sg_float4 *r) {
    const int n = p.nMolMe;
    if (n <= 0) return 0;
    const double Rh = 2.0 * std::cbrt(3.0 * n / (16.0 * M_PI));  // radius in half-lattice units
    const int K = (int)std::ceil(Rh) + 3;
    struct Site {
        long long d2;
        int i, j, k;
    };
    std::vector<Site> sites;
    for (int k = -K; k <= K; ++k)
        for (int j = -K; j <= K; ++j)
            for (int i = -K; i <= K; ++i) {
                if (((i + j + k) & 1) != 0) continue;  // fcc sites in units of a/2
                sites.push_back({(long long)i * i + (long long)j * j + (long long)k * k, i, j, k});
            }
    if ((int)sites.size() < n) return -1;
    std::sort(sites.begin(), sites.end(), [](const Site &a, const Site &b) {
        if (a.d2 != b.d2) return a.d2 < b.d2;
        if (a.k != b.k) return a.k < b.k;
        if (a.j != b.j) return a.j < b.j;
        return a.i < b.i;
    });
    std::vector<std::array<double, 3>> pts(n);
    const double h = 0.5 * (double)p.a;
    for (int q = 0; q < n; ++q) pts[q] = {h * sites[q].i, h * sites[q].j, h * sites[q].k};
    place_cluster(p, gap, pts, r);
    return 0;
}

int icosahedron_shells(int n) {
    int total = 1;
    for (int s = 1; s < 64; ++s) {
        if (total == n) return s - 1;
        if (total > n) return -1;
        total += 10 * s * s + 2;
    }
    return -1;
}

// Mackay icosahedron with nearest-neighbour distance re along the shells: shell s has its 12
// vertices at s*re/2 times the (edge 2) icosahedron vertices, s-1 atoms inside each of the 30
// edges and (s-1)(s-2)/2 inside each of the 20 faces, 10 s^2 + 2 atoms in all.
int gen_icosahedron(const SimParams &p, double gap, sg_float4 *r) {
    const int shells = icosahedron_shells(p.nMolMe);
    if (shells < 0) return -1;
    const double phi = 0.5 * (1.0 + std::sqrt(5.0));
    std::vector<std::array<double, 3>> V;
    for (int s1 = -1; s1 <= 1; s1 += 2)
        for (int s2 = -1; s2 <= 1; s2 += 2) {
            V.push_back({0.0, (double)s1, s2 * phi});
            V.push_back({(double)s1, s2 * phi, 0.0});
            V.push_back({s2 * phi, 0.0, (double)s1});
        }
    auto d2 = [&](int a, int b) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += (V[a][k] - V[b][k]) * (V[a][k] - V[b][k]);
        return s;
    };
    std::vector<std::pair<int, int>> edges;
    std::vector<std::array<int, 3>> faces;
    for (int a = 0; a < 12; ++a)
        for (int b = a + 1; b < 12; ++b)
            if (std::fabs(d2(a, b) - 4.0) < 1e-9) edges.push_back({a, b});
    for (int a = 0; a < 12; ++a)
        for (int b = a + 1; b < 12; ++b)
            for (int c = b + 1; c < 12; ++c)
                if (std::fabs(d2(a, b) - 4.0) < 1e-9 && std::fabs(d2(b, c) - 4.0) < 1e-9 &&
                    std::fabs(d2(a, c) - 4.0) < 1e-9)
                    faces.push_back({a, b, c});
    const double d = (double)p.re;
    std::vector<std::array<double, 3>> pts;
    pts.push_back({0, 0, 0});
    for (int s = 1; s <= shells; ++s) {
        const double sc = 0.5 * s * d;
        auto P = [&](int a) { return std::array<double, 3>{sc * V[a][0], sc * V[a][1], sc * V[a][2]}; };
        for (int a = 0; a < 12; ++a) pts.push_back(P(a));
        for (auto &e : edges) {
            auto A = P(e.first), B = P(e.second);
            for (int t = 1; t < s; ++t) {
                double u = (double)t / s;
                pts.push_back({A[0] + (B[0] - A[0]) * u, A[1] + (B[1] - A[1]) * u, A[2] + (B[2] - A[2]) * u});
            }
        }
        for (auto &f : faces) {
            auto A = P(f[0]), B = P(f[1]), C = P(f[2]);
            for (int u = 1; u < s; ++u)
                for (int v = 1; u + v < s; ++v) {
                    double fu = (double)u / s, fv = (double)v / s;
                    pts.push_back({A[0] + (B[0] - A[0]) * fu + (C[0] - A[0]) * fv,
                                   A[1] + (B[1] - A[1]) * fu + (C[1] - A[1]) * fv,
                                   A[2] + (B[2] - A[2]) * fu + (C[2] - A[2]) * fv});
                }
        }
    }
    if ((int)pts.size() != p.nMolMe) return -1;
    place_cluster(p, gap, pts, r);
    return 0;
}

}  // namespace sg
