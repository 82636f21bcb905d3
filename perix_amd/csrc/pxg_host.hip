// pxg_host.hip -- host-side C++ pieces of the path that feed the device:
// the structured PD point cloud with its face ghosts, RandomIC and the
// explicit step schedule.  Pure host code (no kernels).
#include "pxg_internal.h"
#include <cmath>
#include <random>

namespace {

double node_coord(double lo, int i1, double h)
{
    // Lagrange2DQuad4MeshGenerator.cpp:69-76: Xmin + (i-1)*dx
    return lo + (double)(i1 - 1) * h;
}

} // namespace

extern "C" int pxg_structured_sizes(const pxg_grid *g, int *n, int *nbulk)
{
    PXG_CHECK(g && (g->dim == 2 || g->dim == 3) && g->nx > 0 && g->ny > 0 && (g->dim == 2 || g->nz > 0),
              PXG_ERR_ARG, "pxg_structured_sizes: bad grid");
    long long nb, ng;
    if (g->dim == 2) { nb = (long long)g->nx * g->ny; ng = 2LL * (g->nx + g->ny); }
    else {
        nb = (long long)g->nx * g->ny * g->nz;
        ng = 2LL * ((long long)g->ny * g->nz + (long long)g->nx * g->nz + (long long)g->nx * g->ny);
    }
    PXG_CHECK(nb + ng <= 2147483647LL, PXG_ERR_ARG, "pxg_structured_sizes: too many nodes");
    if (n) *n = (int)(nb + ng);
    if (nbulk) *nbulk = (int)nb;
    return PXG_OK;
}

extern "C" double pxg_horizon_radius(const pxg_grid *g, double factor)
{
    // ReadPDMeshBlock.cpp:17-23 (point spacing) and PDMesh.h:37-41 (1e-6 margin)
    const double spacing = std::fabs(g->xmax - g->xmin) / (double)g->nx;
    return factor * spacing * (1.0 + 1.0e-6);
}

// Structured point cloud: the quad4/hex8 generator's node coordinates
// (Lagrange2DQuad4MeshGenerator.cpp:40-153, Lagrange3DHex8MeshGenerator.cpp:
// 74-130), one PD point per cell centre and one ghost per boundary face,
// half a cell outside the wall (CreatePDMesh.cpp:451-799).  Ghosts follow the
// generator's group order (left, right, bottom, top[, back, front]) and each
// group's face order; a face centre adds its corner coordinates in
// connectivity order and divides by the corner count.
extern "C" int pxg_structured_cloud(const pxg_grid *g, double *coords, double *vols, int *mirror,
                                    int *wall, double *dxyz)
{
    int N = 0, NB = 0;
    int rc = pxg_structured_sizes(g, &N, &NB);
    if (rc) return rc;
    PXG_CHECK(coords && vols && mirror, PXG_ERR_ARG, "pxg_structured_cloud: NULL output");
    const bool is3d = (g->dim == 3);
    const int nx = g->nx, ny = g->ny, nz = is3d ? g->nz : 1;
    const double hx = (g->xmax - g->xmin) / nx;
    const double hy = (g->ymax - g->ymin) / ny;
    const double hz = is3d ? (g->zmax - g->zmin) / nz : 0.0;
    auto X = [&](int i) { return node_coord(g->xmin, i, hx); };
    auto Y = [&](int j) { return node_coord(g->ymin, j, hy); };
    auto Z = [&](int k) { return node_coord(g->zmin, k, hz); };
    double d[3] = {std::fabs(X(2) - X(1)), std::fabs(Y(2) - Y(1)), is3d ? std::fabs(Z(2) - Z(1)) : 0.0};
    if (dxyz) { dxyz[0] = d[0]; dxyz[1] = d[1]; dxyz[2] = d[2]; }
    const double cellVolume = is3d ? d[0] * d[1] * d[2] : d[0] * d[1];
    int id = 0;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double *c = coords + 3 * (size_t)id;
                c[0] = 0.5 * (X(i) + X(i + 1));
                c[1] = 0.5 * (Y(j) + Y(j + 1));
                c[2] = is3d ? 0.5 * (Z(k) + Z(k + 1)) : 0.0;
                vols[id] = cellVolume;
                mirror[id] = 0;
                if (wall) wall[id] = -1;
                id++;
            }
    // mirror lookup of CreatePDMesh.cpp:529-545 reduces to the adjacent cell
    // index when every bulk point's key is its own cell; check that once
    auto cell_key = [&](double x, double y, double z, long &ix, long &iy, long &iz) {
        ix = std::lround((x - 0.5 * d[0]) / d[0]);
        iy = std::lround((y - 0.5 * d[1]) / d[1]);
        iz = is3d ? std::lround((z - 0.5 * d[2]) / d[2]) : 0L;
    };
    for (int b = 0; b < NB; b++) {
        long ix, iy, iz;
        cell_key(coords[3 * (size_t)b], coords[3 * (size_t)b + 1], coords[3 * (size_t)b + 2], ix, iy, iz);
        PXG_CHECK((iz * ny + iy) * (long)nx + ix == b, PXG_ERR_ARG,
                  "pxg_structured_cloud: cell-centre keys are not unique (degenerate grid)");
    }
    static const int cx8[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int cy8[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int cz8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    // hex8 face connectivities (Lagrange3DHex8MeshGenerator.cpp:140-230)
    static const int faces[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                    {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
    const int nwalls = is3d ? 6 : 4;
    for (int w = 0; w < nwalls; w++) {
        const int axis = w / 2;
        const double sign = (w % 2) ? 1.0 : -1.0;
        const double half = 0.5 * d[axis];
        int nfaces, nf;
        if (!is3d) { nfaces = (axis == 0) ? ny : nx; nf = 2; }
        else { nfaces = (axis == 0) ? ny * nz : (axis == 1 ? nx * nz : nx * ny); nf = 4; }
        for (int f = 0; f < nfaces; f++) {
            double px[4], py[4], pz[4] = {0, 0, 0, 0};
            if (!is3d) {
                // quad4 edge connectivities (Lagrange2DQuad4MeshGenerator.cpp:118-147)
                if (w == 0) { px[0] = X(1); py[0] = Y(f + 2); px[1] = X(1); py[1] = Y(f + 1); }
                else if (w == 1) { px[0] = X(nx + 1); py[0] = Y(f + 1); px[1] = X(nx + 1); py[1] = Y(f + 2); }
                else if (w == 2) { px[0] = X(f + 1); py[0] = Y(1); px[1] = X(f + 2); py[1] = Y(1); }
                else { px[0] = X(f + 2); py[0] = Y(ny + 1); px[1] = X(f + 1); py[1] = Y(ny + 1); }
            } else {
                int i, j, k;
                if (axis == 0) { j = f % ny + 1; k = f / ny + 1; i = (w == 0) ? 1 : nx; }
                else if (axis == 1) { i = f % nx + 1; k = f / nx + 1; j = (w == 2) ? 1 : ny; }
                else { i = f % nx + 1; j = f / nx + 1; k = (w == 4) ? 1 : nz; }
                for (int q = 0; q < 4; q++) {
                    const int c = faces[w][q];
                    px[q] = X(i + cx8[c]); py[q] = Y(j + cy8[c]); pz[q] = Z(k + cz8[c]);
                }
            }
            double cx = 0.0, cy = 0.0, cz = 0.0;
            for (int q = 0; q < nf; q++) { cx += px[q]; cy += py[q]; cz += pz[q]; }
            cx /= nf; cy /= nf; cz /= nf;
            double gx = cx, gy = cy, gz = cz, mx = cx, my = cy, mz = cz;
            if (axis == 0) { gx = cx + sign * half; mx = cx - sign * half; }
            else if (axis == 1) { gy = cy + sign * half; my = cy - sign * half; }
            else { gz = cz + sign * half; mz = cz - sign * half; }
            double *c = coords + 3 * (size_t)id;
            c[0] = gx; c[1] = gy; c[2] = is3d ? gz : 0.0;
            vols[id] = cellVolume;
            if (wall) wall[id] = w;
            long ix, iy, iz;
            cell_key(mx, my, mz, ix, iy, iz);
            mirror[id] = (ix >= 0 && ix < nx && iy >= 0 && iy < ny && iz >= 0 && iz < nz)
                             ? (int)((iz * ny + iy) * (long)nx + ix) + 1 : 0;
            id++;
        }
    }
    return (id == N) ? PXG_OK : PXG_ERR_ARG;
}

// RandomIC::apply (RandomIC.cpp:13-42): std::mt19937_64, top 53 bits mapped
// to [min, max); dofs are 1-based as in the reference's IC entries.
extern "C" int pxg_random_ic(uint64_t seed, double lo, double hi, int nnodes, const int *nodes, int ndof,
                             int ndofs_set, const int *dofs, double *values)
{
    PXG_CHECK(lo < hi, PXG_ERR_ARG, "pxg_random_ic: min must be smaller than max");
    PXG_CHECK(nodes && dofs && values && ndof > 0, PXG_ERR_ARG, "pxg_random_ic: bad arguments");
    std::mt19937_64 gen(seed);
    const double inv53 = 1.0 / 9007199254740992.0;
    const double width = hi - lo;
    for (int q = 0; q < nnodes; q++) {
        const size_t base = (size_t)(nodes[q] - 1) * ndof;
        for (int k = 0; k < ndofs_set; k++) {
            const uint64_t bits = gen() >> 11U;
            values[base + dofs[k] - 1] = lo + width * ((double)bits * inv53);
        }
    }
    return PXG_OK;
}

// Step schedule of TimeStepping::solveExplicitDynamics (TimeStepping.cpp:497-510)
extern "C" int pxg_schedule(double dt, double total, int cap, double *dts, double *ts)
{
    const double tEps = 1.0e-12 * total;
    double t = 0.0;
    int step = 0;
    while (t < total - tEps) {
        const double dtA = std::fmin(dt, total - t);
        const double tNew = t + dtA;
        if (step < cap) { if (dts) dts[step] = dtA; if (ts) ts[step] = tNew; }
        step++;
        t = tNew;
    }
    return step;
}
