/*
 * pxo_mesh.c -- oracle (test infrastructure): structured PD point cloud,
 * horizon families, crack cut and CSR pattern, restated from the reference.
 */
#include "pxo.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* a1.  Lagrange2DQuad4MeshGenerator.cpp:40-153 / Lagrange3DHex8MeshGenerator.cpp
 *      :74-130 give the background node coordinates Xmin+(i-1)*dx; the PD
 *      points are built by PDMesh::createPDMesh (CreatePDMesh.cpp:451-799).   */

static double node_coord(double lo, int i1based, double h)
{
    /* Lagrange2DQuad4MeshGenerator.cpp:69-76: Data.Xmin+(i-1)*dx */
    return lo + (double)(i1based - 1) * h;
}

int pxo_mesh_sizes(const pxo_grid *g, int *n_nodes, int *n_bulk)
{
    long long nb, ng;
    if (g->dim == 2) {
        nb = (long long)g->nx * g->ny;
        ng = 2LL * (g->nx + g->ny);               /* left,right,bottom,top */
    } else if (g->dim == 3) {
        nb = (long long)g->nx * g->ny * g->nz;
        ng = 2LL * ((long long)g->ny * g->nz + (long long)g->nx * g->nz
                    + (long long)g->nx * g->ny);
    } else {
        return -1;
    }
    if (nb + ng > 2147483647LL) return -2;
    if (n_bulk) *n_bulk = (int)nb;
    if (n_nodes) *n_nodes = (int)(nb + ng);
    return 0;
}

double pxo_horizon(const pxo_grid *g, double factor)
{
    /* ReadPDMeshBlock.cpp:17-23 (spacing) and PDMesh.h:37-41 (margin) */
    const double spacing = fabs(g->xmax - g->xmin) / (double)g->nx;
    return factor * spacing * (1.0 + 1.0e-6);
}

/* positionKey of CreatePDMesh.cpp:529-540 reduced to the cell index triple */
static void position_cell(double x, double y, double z, const double *d,
                          int is3d, long *ix, long *iy, long *iz)
{
    *ix = lround((x - 0.5 * d[0]) / d[0]);
    *iy = lround((y - 0.5 * d[1]) / d[1]);
    *iz = is3d ? lround((z - 0.5 * d[2]) / d[2]) : 0L;
}

int pxo_mesh_build(const pxo_grid *g, double *coords, double *vols, int *mirror,
                   int *wall, double *dxyz)
{
    int N, NB;
    if (pxo_mesh_sizes(g, &N, &NB)) return -1;
    const int is3d = (g->dim == 3);
    const int nx = g->nx, ny = g->ny, nz = is3d ? g->nz : 1;
    /* generator spacing: (Xmax-Xmin)/Nx (Lagrange2DQuad4MeshGenerator.cpp:40-41) */
    const double hx = (g->xmax - g->xmin) / nx;
    const double hy = (g->ymax - g->ymin) / ny;
    const double hz = is3d ? (g->zmax - g->zmin) / nz : 0.0;
#define XN(i) node_coord(g->xmin, (i), hx)
#define YN(j) node_coord(g->ymin, (j), hy)
#define ZN(k) node_coord(g->zmin, (k), hz)
    /* cell size from the first element (CreatePDMesh.cpp:485-503) */
    double d[3];
    d[0] = fabs(XN(2) - XN(1));
    d[1] = fabs(YN(2) - YN(1));
    d[2] = is3d ? fabs(ZN(2) - ZN(1)) : 0.0;
    if (dxyz) { dxyz[0] = d[0]; dxyz[1] = d[1]; dxyz[2] = d[2]; }
    const double cellVolume = is3d ? d[0] * d[1] * d[2] : d[0] * d[1];

    /* bulk points: cell centre 0.5*(X(i1)+X(i3|i7)) (CreatePDMesh.cpp:509-522) */
    int id = 0;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double *c = coords + 3 * (size_t)id;
                c[0] = 0.5 * (XN(i) + XN(i + 1));
                c[1] = 0.5 * (YN(j) + YN(j + 1));
                c[2] = is3d ? 0.5 * (ZN(k) + ZN(k + 1)) : 0.0;
                vols[id] = cellVolume;
                mirror[id] = 0;
                if (wall) wall[id] = -1;
                id++;
            }
    /* every bulk key must be its own cell index for the mirror lookup below */
    for (int b = 0; b < NB; b++) {
        long ix, iy, iz;
        position_cell(coords[3 * b], coords[3 * b + 1], coords[3 * b + 2], d, is3d,
                      &ix, &iy, &iz);
        const long e = (iz * ny + iy) * (long)nx + ix;
        if (e != b) return -3;
    }

    /* ghosts: boundary groups in generator order (2D: left,right,bottom,top;
     * 3D: + back,front), each face in its connectivity order; ghost at the
     * face centre +/- half a cell (CreatePDMesh.cpp:639-727). The face centre
     * sums the connectivity nodes in their listed order, then divides. */
    const int nwalls = is3d ? 6 : 4;
    for (int w = 0; w < nwalls; w++) {
        const int axis = w / 2;
        const double sign = (w % 2) ? +1.0 : -1.0;
        const double half = 0.5 * d[axis];
        int nfaces, nf;
        if (!is3d) { nfaces = (axis == 0) ? ny : nx; nf = 2; }
        else if (axis == 0) { nfaces = ny * nz; nf = 4; }
        else if (axis == 1) { nfaces = nx * nz; nf = 4; }
        else { nfaces = nx * ny; nf = 4; }
        for (int f = 0; f < nfaces; f++) {
            double px[4], py[4], pz[4];
            if (!is3d) {
                /* Lagrange2DQuad4MeshGenerator.cpp:118-147 */
                if (w == 0) { int j = f + 1;   /* left [i4,i1] */
                    px[0] = XN(1); py[0] = YN(j + 1); px[1] = XN(1); py[1] = YN(j);
                } else if (w == 1) { int j = f + 1;   /* right [i2,i3] */
                    px[0] = XN(nx + 1); py[0] = YN(j); px[1] = XN(nx + 1); py[1] = YN(j + 1);
                } else if (w == 2) { int i = f + 1;   /* bottom [i1,i2] */
                    px[0] = XN(i); py[0] = YN(1); px[1] = XN(i + 1); py[1] = YN(1);
                } else { int i = f + 1;               /* top [i3,i4] */
                    px[0] = XN(i + 1); py[0] = YN(ny + 1); px[1] = XN(i); py[1] = YN(ny + 1);
                }
                pz[0] = pz[1] = 0.0;
            } else {
                /* Lagrange3DHex8MeshGenerator.cpp:140-230 face connectivities */
                int i, j, k;
                if (axis == 0) { j = f % ny + 1; k = f / ny + 1; i = (w == 0) ? 1 : nx; }
                else if (axis == 1) { i = f % nx + 1; k = f / nx + 1; j = (w == 2) ? 1 : ny; }
                else { i = f % nx + 1; j = f / nx + 1; k = (w == 4) ? 1 : nz; }
                /* corner (di,dj,dk) offsets of i1..i8 */
                static const int cx8[8] = {0, 1, 1, 0, 0, 1, 1, 0};
                static const int cy8[8] = {0, 0, 1, 1, 0, 0, 1, 1};
                static const int cz8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
                static const int faces[6][4] = {
                    {0, 4, 7, 3},   /* left   i1 i5 i8 i4 */
                    {1, 2, 6, 5},   /* right  i2 i3 i7 i6 */
                    {0, 1, 5, 4},   /* bottom i1 i2 i6 i5 */
                    {3, 7, 6, 2},   /* top    i4 i8 i7 i3 */
                    {0, 3, 2, 1},   /* back   i1 i4 i3 i2 */
                    {4, 5, 6, 7}};  /* front  i5 i6 i7 i8 */
                for (int q = 0; q < 4; q++) {
                    const int c = faces[w][q];
                    px[q] = XN(i + cx8[c]); py[q] = YN(j + cy8[c]); pz[q] = ZN(k + cz8[c]);
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
            position_cell(mx, my, mz, d, is3d, &ix, &iy, &iz);
            int m = 0;
            if (ix >= 0 && ix < nx && iy >= 0 && iy < ny && iz >= 0 && iz < nz)
                m = (int)((iz * ny + iy) * (long)nx + ix) + 1;
            mirror[id] = m;
            id++;
        }
    }
#undef XN
#undef YN
#undef ZN
    return (id == N) ? 0 : -4;
}

/* ------------------------------------------------------------------------ */
/* a2.  PDMesh::createNeighborNodes (CreatePDMesh.cpp:910-1130).            */

typedef struct { uint64_t key; int val; } hslot;

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33; return x;
}

static uint64_t encode_cell(int ix, int iy, int iz)
{
    /* CreatePDMesh.cpp:969-975 */
    const long bias = 1L << 20;
    return ((uint64_t)((long)ix + bias) << 42) | ((uint64_t)((long)iy + bias) << 21)
           | (uint64_t)((long)iz + bias);
}

static int hget(const hslot *t, uint64_t mask, uint64_t key)
{
    uint64_t h = mix64(key) & mask;
    while (t[h].val >= 0) {
        if (t[h].key == key) return t[h].val;
        h = (h + 1) & mask;
    }
    return -1;
}

long long pxo_families(int n, const double *coords, int is3d, double r,
                       int *off, int *nbr)
{
    if (n <= 0) { if (off) off[0] = 0; return 0; }
    const double r2 = r * r;
    const double r2Tol = 64.0 * DBL_EPSILON * r2;
    const double cellSize = r;
    double minX = DBL_MAX, minY = DBL_MAX, minZ = DBL_MAX;
    for (int i = 0; i < n; i++) {
        minX = fmin(minX, coords[3 * i + 0]);
        minY = fmin(minY, coords[3 * i + 1]);
        minZ = fmin(minZ, coords[3 * i + 2]);
    }
    if (!is3d) minZ = 0.0;
    int *cix = malloc(sizeof(int) * 3 * (size_t)n);
    uint64_t cap = 1; while (cap < 2ULL * (uint64_t)n + 16) cap <<= 1;
    hslot *tab = malloc(sizeof(hslot) * cap);
    for (uint64_t s = 0; s < cap; s++) tab[s].val = -1;
    int *bucket_of = malloc(sizeof(int) * (size_t)n);
    int nbk = 0;
    int *bk_cell = malloc(sizeof(int) * 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        /* CreatePDMesh.cpp:958-964 */
        const int ix = (int)floor((coords[3 * i + 0] - minX) / cellSize);
        const int iy = (int)floor((coords[3 * i + 1] - minY) / cellSize);
        const int iz = is3d ? (int)floor((coords[3 * i + 2] - minZ) / cellSize) : 0;
        cix[3 * i] = ix; cix[3 * i + 1] = iy; cix[3 * i + 2] = iz;
        const uint64_t key = encode_cell(ix, iy, iz);
        uint64_t h = mix64(key) & (cap - 1);
        while (tab[h].val >= 0 && tab[h].key != key) h = (h + 1) & (cap - 1);
        if (tab[h].val < 0) {            /* first appearance -> new bucket */
            tab[h].key = key; tab[h].val = nbk;
            bk_cell[3 * nbk] = ix; bk_cell[3 * nbk + 1] = iy; bk_cell[3 * nbk + 2] = iz;
            nbk++;
        }
        bucket_of[i] = tab[h].val;
    }
    /* bucket member lists, node-id order (CreatePDMesh.cpp:982-995) */
    int *bk_off = calloc((size_t)nbk + 1, sizeof(int));
    for (int i = 0; i < n; i++) bk_off[bucket_of[i] + 1]++;
    for (int b = 0; b < nbk; b++) bk_off[b + 1] += bk_off[b];
    int *bk_nodes = malloc(sizeof(int) * (size_t)n);
    int *cur = malloc(sizeof(int) * (size_t)(nbk > n ? nbk : n));
    for (int b = 0; b < nbk; b++) cur[b] = bk_off[b];
    for (int i = 0; i < n; i++) bk_nodes[cur[bucket_of[i]]++] = i;

    static const int st2[4][3] = {{1, -1, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
    static const int st3[13][3] = {
        {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0}, {1, 0, 1},
        {1, 1, -1}, {1, 1, 0}, {1, 1, 1}, {0, 1, -1}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}};
    const int (*st)[3] = is3d ? st3 : st2;
    const int nst = is3d ? 13 : 4;

    /* pass 1 counts, pass 2 fills with per-node cursors */
    const int fill = (nbr != NULL);
    int *cnt = NULL;
    if (!fill) { memset(off, 0, sizeof(int) * ((size_t)n + 1)); }
    else { cnt = malloc(sizeof(int) * (size_t)n); for (int i = 0; i < n; i++) cnt[i] = off[i]; }
    long long nb = 0;
#define TRY_ADD(I, J) do { \
        const int ii_ = (I), jj_ = (J); \
        const double dxij = coords[3 * ii_] - coords[3 * jj_]; \
        const double dyij = coords[3 * ii_ + 1] - coords[3 * jj_ + 1]; \
        const double dzij = is3d ? (coords[3 * ii_ + 2] - coords[3 * jj_ + 2]) : 0.0; \
        const double dist2 = dxij * dxij + dyij * dyij + dzij * dzij; \
        if (dist2 <= r2 + r2Tol) { \
            if (fill) { nbr[cnt[ii_]++] = jj_ + 1; nbr[cnt[jj_]++] = ii_ + 1; } \
            else { off[ii_ + 1]++; off[jj_ + 1]++; } \
            nb += 2; } } while (0)
    for (int b = 0; b < nbk; b++) {
        const int b0 = bk_off[b], b1 = bk_off[b + 1];
        for (int li = b0; li < b1; li++)
            for (int lj = li + 1; lj < b1; lj++) TRY_ADD(bk_nodes[li], bk_nodes[lj]);
        for (int o = 0; o < nst; o++) {
            const int nbk_id = hget(tab, cap - 1,
                                    encode_cell(bk_cell[3 * b] + st[o][0],
                                                bk_cell[3 * b + 1] + st[o][1],
                                                bk_cell[3 * b + 2] + st[o][2]));
            if (nbk_id < 0) continue;
            const int c0 = bk_off[nbk_id], c1 = bk_off[nbk_id + 1];
            for (int li = b0; li < b1; li++)
                for (int lj = c0; lj < c1; lj++) TRY_ADD(bk_nodes[li], bk_nodes[lj]);
        }
    }
#undef TRY_ADD
    if (!fill) for (int i = 0; i < n; i++) off[i + 1] += off[i];
    free(cnt); free(cix); free(tab); free(bucket_of); free(bk_cell);
    free(bk_off); free(bk_nodes); free(cur);
    return nb;
}

/* ------------------------------------------------------------------------ */
/* a3.  initialCrackCutsBond (PDMesh.cpp:104-159) with orientSide (:30-37),
 *      planeSide (:45-54) and bondPiercesQuad (:67-98).                    */

static int orient_side(double ux, double uy, double vx, double vy)
{
    const double det = ux * vy - uy * vx;
    const double scale = fabs(ux) * fabs(vy) + fabs(uy) * fabs(vx);
    if (det > 1.0e-12 * scale) return +1;
    if (det < -1.0e-12 * scale) return -1;
    return 0;
}

static int plane_side(double px, double py, double pz, double c0x, double c0y,
                      double c0z, double nx, double ny, double nz)
{
    const double wx = px - c0x, wy = py - c0y, wz = pz - c0z;
    const double d = wx * nx + wy * ny + wz * nz;
    const double scale = fabs(wx * nx) + fabs(wy * ny) + fabs(wz * nz);
    if (d > 1.0e-12 * scale) return +1;
    if (d < -1.0e-12 * scale) return -1;
    return 0;
}

static int bond_pierces_quad(double ax, double ay, double az, double bx,
                             double by, double bz, const double *q)
{
    const double cx[4] = {q[0], q[3], q[6], q[9]};
    const double cy[4] = {q[1], q[4], q[7], q[10]};
    const double cz[4] = {q[2], q[5], q[8], q[11]};
    const double ux = cx[1] - cx[0], uy = cy[1] - cy[0], uz = cz[1] - cz[0];
    const double vx = cx[3] - cx[0], vy = cy[3] - cy[0], vz = cz[3] - cz[0];
    const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    const double nlen = sqrt(nx * nx + ny * ny + nz * nz);
    if (nlen <= 0.0) return 0;
    const int sA = plane_side(ax, ay, az, cx[0], cy[0], cz[0], nx, ny, nz);
    const int sB = plane_side(bx, by, bz, cx[0], cy[0], cz[0], nx, ny, nz);
    if (sA * sB >= 0) return 0;
    const double dA = (ax - cx[0]) * nx + (ay - cy[0]) * ny + (az - cz[0]) * nz;
    const double dB = (bx - cx[0]) * nx + (by - cy[0]) * ny + (bz - cz[0]) * nz;
    const double t = dA / (dA - dB);
    const double Qx = ax + t * (bx - ax), Qy = ay + t * (by - ay), Qz = az + t * (bz - az);
    const double unx = nx / nlen, uny = ny / nlen, unz = nz / nlen;
    const double tol = 1.0e-9 * nlen;
    for (int k = 0; k < 4; k++) {
        const int m = (k + 1) % 4;
        const double ex = cx[m] - cx[k], ey = cy[m] - cy[k], ez = cz[m] - cz[k];
        const double gx = Qx - cx[k], gy = Qy - cy[k], gz = Qz - cz[k];
        const double wx = ey * gz - ez * gy, wy = ez * gx - ex * gz, wz = ex * gy - ey * gx;
        if (wx * unx + wy * uny + wz * unz <= tol) return 0;
    }
    return 1;
}

int pxo_crack_cuts_bond(const pxo_cracks *c, int n, const double *coords,
                        const int *mirror, int i1, int j1)
{
    if (!c || (c->nseg == 0 && c->nquad == 0)) return 0;
    int a = i1, b = j1;
    if (mirror) {
        if (a >= 1 && a <= n && mirror[a - 1] >= 1 && mirror[a - 1] <= n) a = mirror[a - 1];
        if (b >= 1 && b <= n && mirror[b - 1] >= 1 && mirror[b - 1] <= n) b = mirror[b - 1];
    }
    if (a == b) return 0;
    const double ax = coords[3 * (a - 1)], ay = coords[3 * (a - 1) + 1], az = coords[3 * (a - 1) + 2];
    const double bx = coords[3 * (b - 1)], by = coords[3 * (b - 1) + 1], bz = coords[3 * (b - 1) + 2];
    for (int s = 0; s < c->nseg; s++) {
        const double *k = c->seg + 4 * s;
        const double cx = k[0], cy = k[1], dx = k[2], dy = k[3];
        const int s1 = orient_side(dx - cx, dy - cy, ax - cx, ay - cy);
        const int s2 = orient_side(dx - cx, dy - cy, bx - cx, by - cy);
        if (s1 * s2 >= 0) continue;
        const int s3 = orient_side(bx - ax, by - ay, cx - ax, cy - ay);
        const int s4 = orient_side(bx - ax, by - ay, dx - ax, dy - ay);
        if (s3 * s4 < 0) return 1;
    }
    for (int q = 0; q < c->nquad; q++)
        if (bond_pierces_quad(ax, ay, az, bx, by, bz, c->quad + 12 * q)) return 1;
    return 0;
}

void pxo_seed_health(const pxo_cracks *c, int n, const double *coords,
                     const int *mirror, const int *off, const int *nbr,
                     double *health)
{
    /* ExplicitPDDOFracElement.cpp:134-155 (force-only seeding) */
    for (int i = 0; i < n; i++)
        for (int t = off[i]; t < off[i + 1]; t++)
            health[t] = pxo_crack_cuts_bond(c, n, coords, mirror, i + 1, nbr[t]) ? 0.0 : 1.0;
}

/* ------------------------------------------------------------------------ */
/* a4.  SparseMatrix::initFromPDMesh + setup (SparseMatrix.cpp:73-185).     */

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

long long pxo_csr_pattern(int n, int ndof, const int *off, const int *nbr,
                          const int *mirror, int *rowptr, int *cols)
{
    int maxf = 0;
    for (int i = 0; i < n; i++) if (off[i + 1] - off[i] > maxf) maxf = off[i + 1] - off[i];
    int *tmp = malloc(sizeof(int) * ((size_t)maxf + 2));
    long long nnz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int i = 0; i < n; i++) {
        int m = 0;
        tmp[m++] = i + 1;
        for (int t = off[i]; t < off[i + 1]; t++) tmp[m++] = nbr[t];
        if (mirror && mirror[i] >= 1 && mirror[i] <= n) tmp[m++] = mirror[i];
        qsort(tmp, (size_t)m, sizeof(int), cmp_int);
        int u = 0;
        for (int k = 0; k < m; k++) if (u == 0 || tmp[k] != tmp[u - 1]) tmp[u++] = tmp[k];
        for (int a = 0; a < ndof; a++) {
            const long long row = (long long)i * ndof + a;
            if (cols) {
                long long p = rowptr[row];
                for (int k = 0; k < u; k++)
                    for (int b = 0; b < ndof; b++) cols[p++] = (tmp[k] - 1) * ndof + b;
            }
            nnz += (long long)u * ndof;
            if (rowptr) rowptr[row + 1] = (int)nnz;
        }
    }
    free(tmp);
    return nnz;
}
