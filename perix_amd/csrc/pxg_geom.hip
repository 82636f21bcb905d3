// pxg_geom.hip -- point cloud upload, horizon families, crack seeds and the
// implicit CSR pattern, all built on the device.
//
// Families (reference: PDMesh::createNeighborNodes, CreatePDMesh.cpp:910-1130).
// The reference sweeps cell buckets (side r, first-appearance order) and,
// per bucket, its intra pairs then a forward half-stencil, pushing every
// accepted pair into both nodes' lists.  Node k therefore meets its
// neighbours bucket by bucket: when its own bucket B is processed it gets
// B's other members, then the members of B+s_o for each forward offset o;
// when a bucket B' = B - s_o is processed it gets B''s members.  Buckets are
// processed in order of their smallest node id.  So each node can rebuild
// its own list independently: sort its 3^d neighbour cells by the key
// (first node id of the processing bucket, 0 | 1+o) and append the members
// of each cell (ascending id) that pass the distance test.  No atomics touch
// the lists, and the order is the reference's by construction.
#include "pxg_internal.h"
#include <algorithm>
#include <cfloat>
#include <cmath>

namespace pxg {

// ---------------------------------------------------------------- scan --
// exclusive scan of int counts into int offsets (64-bit total, checked)
__global__ void k_scan_blocks(const int *in, long long n, long long *out, long long *bsum)
{
    __shared__ long long s[1024];
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    long long v = (i < n) ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        long long t = (threadIdx.x >= (unsigned)d) ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}

__global__ void k_scan_bsum(long long *bsum, long long nb, long long *total)
{
    // single thread: nb is at most a few thousand
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        long long acc = 0;
        for (long long b = 0; b < nb; b++) { long long v = bsum[b]; bsum[b] = acc; acc += v; }
        *total = acc;
    }
}

__global__ void k_scan_add(const long long *tmp, const long long *bsum, long long n, int *out)
{
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] = (int)(tmp[i] + bsum[blockIdx.x]);
}

// out[0..n] (n+1 entries): out[n] = total
int scan_counts(hipStream_t st, const int *counts, long long n, int *out, long long *total_host)
{
    const long long nblk = (n + 1023) / 1024 + 1;
    long long *tmp = nullptr, *bsum = nullptr;
    PXG_HIP(hipMalloc(&tmp, sizeof(long long) * (n + 1)));
    PXG_HIP(hipMalloc(&bsum, sizeof(long long) * (nblk + 1)));
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)((n + 1023) / 1024 + 1)), dim3(1024), 0, st,
                       counts, n, tmp, bsum);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1), 0, st, bsum, (n + 1023) / 1024 + 1, bsum + nblk);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 1023) / 1024 + 1)), dim3(1024), 0, st,
                       tmp, bsum, n, out);
    long long total = 0;
    PXG_HIP(hipMemcpyAsync(&total, bsum + nblk, sizeof(long long), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    PXG_HIP(hipFree(tmp));
    PXG_HIP(hipFree(bsum));
    PXG_CHECK(total <= 2147483647LL, PXG_ERR_ARG,
              "more than 2^31-1 entries: the 32-bit offsets would overflow "
              "(same guard as ExplicitPDDOFracElement.cpp:116-130 / SparseMatrix.cpp:137-155)");
    int t32 = (int)total;
    PXG_HIP(hipMemcpyAsync(out + n, &t32, sizeof(int), hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    *total_host = total;
    return PXG_OK;
}

// ------------------------------------------------------------- points --
__global__ void k_pack_coords(const double *c3, int n, int dim, double *xp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < dim; a++) xp[(size_t)i * dim + a] = c3[3 * (size_t)i + a];
}

// Bounding box of the points in two passes: 256-thread blocks reduce
// block-strided slices to partials part[6*b + {minx,miny,minz,maxx,maxy,maxz}],
// then one block folds the partials.  min and max are exact, so the split
// does not change the result.
constexpr int kMinmaxThreads = 256, kMinmaxParts = 256;

__device__ void minmax_block(double (*s)[kMinmaxThreads], double *v)
{
    for (int a = 0; a < 6; a++) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int d = kMinmaxThreads / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int a = 0; a < 6; a++)
                s[a][threadIdx.x] = (a < 3) ? fmin(s[a][threadIdx.x], s[a][threadIdx.x + d])
                                            : fmax(s[a][threadIdx.x], s[a][threadIdx.x + d]);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kMinmaxThreads) k_minmax_part(const double *c3, int n, double *part)
{
    __shared__ double s[6][kMinmaxThreads];
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (long long i = (long long)blockIdx.x * kMinmaxThreads + threadIdx.x; i < n;
         i += (long long)gridDim.x * kMinmaxThreads)
        for (int a = 0; a < 3; a++) {
            const double x = c3[3 * (size_t)i + a];
            v[a] = fmin(v[a], x);
            v[3 + a] = fmax(v[3 + a], x);
        }
    minmax_block(s, v);
    if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(kMinmaxThreads) k_minmax_fold(const double *part, int nparts, double *out)
{
    __shared__ double s[6][kMinmaxThreads];
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int p = threadIdx.x; p < nparts; p += kMinmaxThreads)
        for (int a = 0; a < 6; a++)
            v[a] = (a < 3) ? fmin(v[a], part[6 * (size_t)p + a]) : fmax(v[a], part[6 * (size_t)p + a]);
    minmax_block(s, v);
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ---------------------------------------------------------- cell list --
struct CellGrid {
    double minX, minY, minZ, r;
    int ncx, ncy, ncz, is3d;
};

__device__ __forceinline__ void cell_ijk(const double *xp, int dim, int i, const CellGrid &g,
                                         int &ix, int &iy, int &iz)
{
    // CreatePDMesh.cpp:958-964
    ix = (int)floor((xp[(size_t)i * dim + 0] - g.minX) / g.r);
    iy = (int)floor((xp[(size_t)i * dim + 1] - g.minY) / g.r);
    iz = g.is3d ? (int)floor((xp[(size_t)i * dim + 2] - g.minZ) / g.r) : 0;
}

__global__ void k_cell_count(const double *xp, int n, int dim, CellGrid g, int *cell, int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix, iy, iz;
    cell_ijk(xp, dim, i, g, ix, iy, iz);
    const int c = (iz * g.ncy + iy) * g.ncx + ix;
    cell[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void k_cell_fill(const int *cell, int n, const int *start, int *cursor, int *members)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell[i];
    const int p = atomicAdd(&cursor[c], 1);
    members[start[c] + p] = i;
}

__global__ void k_cell_sort(const int *start, int ncell, int *members)
{
    // members of one cell ascending by node id (the reference appends nodes
    // to their bucket in id order, CreatePDMesh.cpp:982-995)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = start[c], e = start[c + 1];
    for (int a = b + 1; a < e; a++) {
        const int v = members[a];
        int k = a - 1;
        while (k >= b && members[k] > v) { members[k + 1] = members[k]; k--; }
        members[k + 1] = v;
    }
}

// forward half-stencils (CreatePDMesh.cpp:1001-1010); class of offset
// (dx,dy,dz): -2 self, 0..12 forward index, -1 backward
__constant__ signed char c_class3[27];
__constant__ signed char c_class2[9];

template <bool FILL>
__global__ void k_families(const double *xp, int n, int dim, CellGrid g, const int *start,
                           const int *members, const int *cell, double thr, int *count_or_off,
                           int *nbr)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int ix, iy, iz;
    cell_ijk(xp, dim, k, g, ix, iy, iz);
    const int ck = cell[k];
    const long long minc = members[start[ck]];
    long long key[27];
    int cid[27];
    int ne = 0;
    const int zr = g.is3d ? 1 : 0;
    for (int dz = -zr; dz <= zr; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int cx = ix + dx, cy = iy + dy, cz = iz + dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= g.ncx || cy >= g.ncy || cz >= g.ncz) continue;
                const int c = (cz * g.ncy + cy) * g.ncx + cx;
                if (start[c + 1] == start[c]) continue;
                const int cls = g.is3d ? c_class3[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]
                                       : c_class2[(dy + 1) * 3 + (dx + 1)];
                long long kk;
                if (cls == -2) kk = minc << 5;                          // own bucket: intra pairs
                else if (cls >= 0) kk = (minc << 5) | (1 + cls);        // forward offset of own bucket
                else kk = ((long long)members[start[c]]) << 5;          // processed by that bucket
                int p = ne++;
                while (p > 0 && key[p - 1] > kk) { key[p] = key[p - 1]; cid[p] = cid[p - 1]; p--; }
                key[p] = kk; cid[p] = c;
            }
    const int dimv = dim;
    const double xk = xp[(size_t)k * dimv], yk = xp[(size_t)k * dimv + 1];
    const double zk = g.is3d ? xp[(size_t)k * dimv + 2] : 0.0;
    int cnt = 0;
    int pos = FILL ? count_or_off[k] : 0;
    for (int e = 0; e < ne; e++) {
        const int c = cid[e];
        for (int q = start[c]; q < start[c + 1]; q++) {
            const int m = members[q];
            if (m == k) continue;
            const double dxij = xk - xp[(size_t)m * dimv];
            const double dyij = yk - xp[(size_t)m * dimv + 1];
            const double dzij = g.is3d ? (zk - xp[(size_t)m * dimv + 2]) : 0.0;
            const double dist2 = dxij * dxij + dyij * dyij + dzij * dzij;
            if (dist2 <= thr) {
                if (FILL) nbr[pos++] = m + 1;
                else cnt++;
            }
        }
    }
    if (!FILL) count_or_off[k] = cnt;
}

// ------------------------------------------------------------- cracks --
// initialCrackCutsBond (PDMesh.cpp:104-159), orientSide (:30-37),
// planeSide (:45-54), bondPiercesQuad (:67-98)
__device__ int orient_side(double ux, double uy, double vx, double vy)
{
    const double det = ux * vy - uy * vx;
    const double scale = fabs(ux) * fabs(vy) + fabs(uy) * fabs(vx);
    if (det > 1.0e-12 * scale) return +1;
    if (det < -1.0e-12 * scale) return -1;
    return 0;
}

__device__ int plane_side(double px, double py, double pz, double c0x, double c0y, double c0z,
                          double nx, double ny, double nz)
{
    const double wx = px - c0x, wy = py - c0y, wz = pz - c0z;
    const double d = wx * nx + wy * ny + wz * nz;
    const double scale = fabs(wx * nx) + fabs(wy * ny) + fabs(wz * nz);
    if (d > 1.0e-12 * scale) return +1;
    if (d < -1.0e-12 * scale) return -1;
    return 0;
}

__device__ int pierces_quad(double ax, double ay, double az, double bx, double by, double bz,
                            const double *q)
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

__global__ void k_seed(const double *c3, const int *mirror, int n, const int *off, const int *nbr,
                       const double *seg, int nseg, const double *quad, int nquad,
                       unsigned char *seed)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    int a = i + 1;
    if (mirror[i] >= 1 && mirror[i] <= n) a = mirror[i];
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) {
        int b = nbr[t];
        if (b >= 1 && b <= n && mirror[b - 1] >= 1 && mirror[b - 1] <= n) b = mirror[b - 1];
        unsigned char cut = 0;
        if (a != b) {
            const double ax = c3[3 * (size_t)(a - 1)], ay = c3[3 * (size_t)(a - 1) + 1],
                         az = c3[3 * (size_t)(a - 1) + 2];
            const double bx = c3[3 * (size_t)(b - 1)], by = c3[3 * (size_t)(b - 1) + 1],
                         bz = c3[3 * (size_t)(b - 1) + 2];
            for (int s = 0; s < nseg && !cut; s++) {
                const double *k = seg + 4 * s;
                const double cx = k[0], cy = k[1], dx = k[2], dy = k[3];
                const int s1 = orient_side(dx - cx, dy - cy, ax - cx, ay - cy);
                const int s2 = orient_side(dx - cx, dy - cy, bx - cx, by - cy);
                if (s1 * s2 >= 0) continue;
                const int s3 = orient_side(bx - ax, by - ay, cx - ax, cy - ay);
                const int s4 = orient_side(bx - ax, by - ay, dx - ax, dy - ay);
                if (s3 * s4 < 0) cut = 1;
            }
            for (int q = 0; q < nquad && !cut; q++)
                if (pierces_quad(ax, ay, az, bx, by, bz, quad + 12 * q)) cut = 1;
        }
        seed[t] = cut ? 0 : 1;
    }
}

// ---------------------------------------------------------------- CSR --
// SparseMatrix::initFromPDMesh (SparseMatrix.cpp:73-115): row block i holds
// {i} U H_i U {mirror(i)}, sorted and unique, each expanded to ndof columns.
constexpr int kMaxRow = 512;

__device__ int row_nodes(int i, const int *off, const int *nbr, const int *mirror, int n, int *tmp)
{
    int m = 0;
    tmp[m++] = i + 1;
    for (int t = off[i]; t < off[i + 1] && m < kMaxRow - 1; t++) tmp[m++] = nbr[t];
    if (mirror[i] >= 1 && mirror[i] <= n) tmp[m++] = mirror[i];
    for (int a = 1; a < m; a++) {
        const int v = tmp[a];
        int k = a - 1;
        while (k >= 0 && tmp[k] > v) { tmp[k + 1] = tmp[k]; k--; }
        tmp[k + 1] = v;
    }
    int u = 0;
    for (int k = 0; k < m; k++) if (u == 0 || tmp[k] != tmp[u - 1]) tmp[u++] = tmp[k];
    return u;
}

__global__ void k_csr_count(const int *off, const int *nbr, const int *mirror, int n, int ndof,
                            int *rowcount, int *flags)
{
    // flags[0]: a family exceeds the row limit; flags[1]: most blocks in a row
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (off[i + 1] - off[i] > kMaxRow - 2) { atomicExch(flags, 1); return; }
    int tmp[kMaxRow];
    const int u = row_nodes(i, off, nbr, mirror, n, tmp);
    for (int a = 0; a < ndof; a++) rowcount[(size_t)i * ndof + a] = u * ndof;
    atomicMax(flags + 1, u);
}

__global__ void k_csr_fill(const int *off, const int *nbr, const int *mirror, int n, int ndof,
                           const int *rowptr, int *cols, int *slot, int *diag, int *mslot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int tmp[kMaxRow];
    const int u = row_nodes(i, off, nbr, mirror, n, tmp);
    for (int a = 0; a < ndof; a++) {
        int p = rowptr[(size_t)i * ndof + a];
        for (int k = 0; k < u; k++)
            for (int b = 0; b < ndof; b++) cols[p++] = (tmp[k] - 1) * ndof + b;
    }
    const int r0 = rowptr[(size_t)i * ndof];
    auto find = [&](int node1) {
        int lo = 0, hi = u;
        while (lo < hi) { const int mid = (lo + hi) >> 1; if (tmp[mid] < node1) lo = mid + 1; else hi = mid; }
        return lo;
    };
    for (int t = off[i]; t < off[i + 1]; t++) slot[t] = r0 + find(nbr[t]) * ndof;
    diag[i] = r0 + find(i + 1) * ndof;
    int ms = -1;
    const int mi = mirror[i];
    if (mi >= 1 && mi <= n && mi != i + 1) {
        bool infam = false;
        for (int t = off[i]; t < off[i + 1]; t++) if (nbr[t] == mi) { infam = true; break; }
        if (!infam) ms = r0 + find(mi) * ndof;
    }
    mslot[i] = ms;
}

} // namespace pxg

using namespace pxg;

static unsigned grid_of(long long n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" int pxg_set_points(pxg_ctx *ctx, int n, int dim, const double *coords,
                              const double *vols, const int *mirror, const double *dxyz)
{
    PXG_CHECK(ctx && n > 0 && (dim == 2 || dim == 3) && coords && vols && mirror && dxyz,
              PXG_ERR_ARG, "pxg_set_points: bad arguments");
    PXG_FAULT(ctx, "pxg_set_points");
    PXG_HIP(hipSetDevice(ctx->device));
    ctx->n = n; ctx->dim = dim;
    for (int a = 0; a < 3; a++) ctx->dxyz[a] = dxyz[a];
    ctx->hcoords.assign(coords, coords + 3 * (size_t)n);
    ctx->hmirror.assign(mirror, mirror + n);
    PXG_HIP(ctx->coords.alloc((size_t)3 * n));
    PXG_HIP(ctx->xp.alloc((size_t)dim * n));
    PXG_HIP(ctx->vols.alloc(n));
    PXG_HIP(ctx->mirror.alloc(n));
    PXG_HIP(hipMemcpyAsync(ctx->coords.p, coords, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    PXG_HIP(hipMemcpyAsync(ctx->vols.p, vols, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    PXG_HIP(hipMemcpyAsync(ctx->mirror.p, mirror, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_pack_coords, dim3(grid_of(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->coords.p, n, dim, ctx->xp.p);
    PXG_HIP(hipGetLastError());
    // invalidate everything derived from the old cloud
    ctx->nb = 0; ctx->nop = 0; ctx->nnz = 0;
    invalidate(ctx, INV_OPS | INV_CRACKS);
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_build_families(pxg_ctx *ctx, double r, long long *nb_out)
{
    PXG_CHECK(ctx && ctx->n > 0, PXG_ERR_STATE, "pxg_build_families: no points");
    PXG_CHECK(r > 0.0, PXG_ERR_ARG, "pxg_build_families: horizon radius must be positive");
    PXG_FAULT(ctx, "pxg_build_families");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int n = ctx->n, dim = ctx->dim;
    dbuf<double> mm;
    const int nparts = (int)std::min<long long>(kMinmaxParts, ((long long)n + kMinmaxThreads - 1) / kMinmaxThreads);
    PXG_HIP(mm.alloc(6 + 6 * (size_t)nparts));
    hipLaunchKernelGGL(k_minmax_part, dim3(nparts), dim3(kMinmaxThreads), 0, st, ctx->coords.p, n, mm.p + 6);
    hipLaunchKernelGGL(k_minmax_fold, dim3(1), dim3(kMinmaxThreads), 0, st, mm.p + 6, nparts, mm.p);
    double h[6];
    PXG_HIP(hipMemcpyAsync(h, mm.p, sizeof(h), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    mm.release();
    CellGrid g;
    g.is3d = (dim == 3);
    g.minX = h[0]; g.minY = h[1]; g.minZ = g.is3d ? h[2] : 0.0; g.r = r;
    const double ex = floor((h[3] - g.minX) / r), ey = floor((h[4] - g.minY) / r),
                 ez = g.is3d ? floor((h[5] - g.minZ) / r) : 0.0;
    PXG_CHECK(ex < 2e6 && ey < 2e6 && ez < 2e6 && (ex + 1) * (ey + 1) * (ez + 1) < 1.5e9,
              PXG_ERR_ARG, "pxg_build_families: horizon too small for the cell grid");
    g.ncx = (int)ex + 1; g.ncy = (int)ey + 1; g.ncz = (int)ez + 1;
    const long long ncell = (long long)g.ncx * g.ncy * g.ncz;
    // stencil classes
    {
        signed char c3[27], c2[9];
        static const int st3[13][3] = {{1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},
                                       {1, 0, 1}, {1, 1, -1}, {1, 1, 0}, {1, 1, 1}, {0, 1, -1},
                                       {0, 1, 0}, {0, 1, 1}, {0, 0, 1}};
        static const int st2[4][2] = {{1, -1}, {1, 0}, {1, 1}, {0, 1}};
        for (int i = 0; i < 27; i++) c3[i] = -1;
        for (int i = 0; i < 9; i++) c2[i] = -1;
        c3[13] = -2; c2[4] = -2;
        for (int o = 0; o < 13; o++) c3[(st3[o][2] + 1) * 9 + (st3[o][1] + 1) * 3 + st3[o][0] + 1] = (signed char)o;
        for (int o = 0; o < 4; o++) c2[(st2[o][1] + 1) * 3 + st2[o][0] + 1] = (signed char)o;
        PXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_class3), c3, sizeof(c3), 0, hipMemcpyHostToDevice));
        PXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_class2), c2, sizeof(c2), 0, hipMemcpyHostToDevice));
    }
    dbuf<int> cell, count, start, cursor, members, famc;
    PXG_HIP(cell.alloc(n));
    PXG_HIP(count.alloc(ncell + 1));
    PXG_HIP(start.alloc(ncell + 1));
    PXG_HIP(cursor.alloc(ncell + 1));
    PXG_HIP(members.alloc(n));
    PXG_HIP(hipMemsetAsync(count.p, 0, sizeof(int) * (ncell + 1), st));
    PXG_HIP(hipMemsetAsync(cursor.p, 0, sizeof(int) * (ncell + 1), st));
    hipLaunchKernelGGL(k_cell_count, dim3(grid_of(n, 256)), dim3(256), 0, st, ctx->xp.p, n, dim, g,
                       cell.p, count.p);
    long long tot = 0;
    int rc = scan_counts(st, count.p, ncell, start.p, &tot);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cell_fill, dim3(grid_of(n, 256)), dim3(256), 0, st, cell.p, n, start.p,
                       cursor.p, members.p);
    hipLaunchKernelGGL(k_cell_sort, dim3(grid_of(ncell, 256)), dim3(256), 0, st, start.p, (int)ncell,
                       members.p);
    const double r2 = r * r;
    const double thr = r2 + 64.0 * DBL_EPSILON * r2;   // CreatePDMesh.cpp:924-926
    PXG_HIP(famc.alloc(n));
    hipLaunchKernelGGL(k_families<false>, dim3(grid_of(n, 128)), dim3(128), 0, st, ctx->xp.p, n, dim,
                       g, start.p, members.p, cell.p, thr, famc.p, (int *)nullptr);
    PXG_HIP(ctx->off.alloc((size_t)n + 1));
    rc = scan_counts(st, famc.p, n, ctx->off.p, &tot);
    if (rc) return rc;
    ctx->nb = tot;
    PXG_HIP(ctx->nbr.alloc((size_t)(tot > 0 ? tot : 1)));
    hipLaunchKernelGGL(k_families<true>, dim3(grid_of(n, 128)), dim3(128), 0, st, ctx->xp.p, n, dim,
                       g, start.p, members.p, cell.p, thr, ctx->off.p, ctx->nbr.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(st));
    cell.release(); count.release(); start.release(); cursor.release(); members.release(); famc.release();
    ctx->delta = r;
    ctx->nop = 0; ctx->nnz = 0;
    invalidate(ctx, INV_OPS | INV_CRACKS);
    // crack seeds follow the families
    PXG_HIP(ctx->seed.alloc((size_t)(tot > 0 ? tot : 1)));
    if (ctx->nseg + ctx->nquad > 0) {
        hipLaunchKernelGGL(k_seed, dim3(grid_of(n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->coords.p,
                           ctx->mirror.p, n, ctx->off.p, ctx->nbr.p, ctx->crack_seg.p, ctx->nseg,
                           ctx->crack_quad.p, ctx->nquad, ctx->seed.p);
    } else {
        PXG_HIP(hipMemsetAsync(ctx->seed.p, 1, (size_t)(tot > 0 ? tot : 1), st));
    }
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(st));
    if (nb_out) *nb_out = tot;
    return PXG_OK;
}

// seed of a new family set: the crack cut when cracks are set, else all intact
static int reseed(pxg_ctx *ctx)
{
    const long long nb = ctx->nb;
    PXG_HIP(ctx->seed.alloc((size_t)(nb > 0 ? nb : 1)));
    if (ctx->nseg + ctx->nquad > 0 && nb > 0)
        hipLaunchKernelGGL(k_seed, dim3(grid_of(ctx->n, kConvNodes)), dim3(kConvBlock), 0, ctx->stream, ctx->coords.p,
                           ctx->mirror.p, ctx->n, ctx->off.p, ctx->nbr.p, ctx->crack_seg.p, ctx->nseg,
                           ctx->crack_quad.p, ctx->nquad, ctx->seed.p);
    else
        PXG_HIP(hipMemsetAsync(ctx->seed.p, 1, (size_t)(nb > 0 ? nb : 1), ctx->stream));
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_set_families(pxg_ctx *ctx, const int *off, const int *nbr)
{
    PXG_CHECK(ctx && ctx->n > 0, PXG_ERR_STATE, "pxg_set_families: no points");
    PXG_CHECK(off, PXG_ERR_ARG, "pxg_set_families: bad arguments");
    const int n = ctx->n;
    PXG_CHECK(off[0] == 0, PXG_ERR_ARG, "pxg_set_families: off[0] must be 0");
    for (int i = 0; i < n; i++)
        PXG_CHECK(off[i + 1] >= off[i], PXG_ERR_ARG, "pxg_set_families: offsets must not decrease");
    const long long nb = off[n];
    PXG_CHECK(nb == 0 || nbr, PXG_ERR_ARG, "pxg_set_families: bad arguments");
    for (long long t = 0; t < nb; t++)
        PXG_CHECK(nbr[t] >= 1 && nbr[t] <= n, PXG_ERR_ARG,
                  "pxg_set_families: neighbour id outside 1..n (ids are 1-based)");
    PXG_FAULT(ctx, "pxg_set_families");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PXG_HIP(ctx->off.alloc((size_t)n + 1));
    PXG_HIP(ctx->nbr.alloc((size_t)(nb > 0 ? nb : 1)));
    PXG_HIP(hipMemcpyAsync(ctx->off.p, off, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, st));
    if (nb > 0) PXG_HIP(hipMemcpyAsync(ctx->nbr.p, nbr, sizeof(int) * (size_t)nb, hipMemcpyHostToDevice, st));
    ctx->nb = nb;
    ctx->nop = 0; ctx->nnz = 0;
    invalidate(ctx, INV_OPS | INV_CRACKS);
    const int rc = reseed(ctx);
    if (rc) return rc;
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_set_csr(pxg_ctx *ctx, int ndof, const int *rowptr, const int *cols)
{
    PXG_CHECK(ctx && ctx->nb > 0, PXG_ERR_STATE, "pxg_set_csr: no families");
    PXG_CHECK(ndof >= 1 && ndof <= 8 && rowptr && cols, PXG_ERR_ARG, "pxg_set_csr: bad arguments");
    const int n = ctx->n;
    const long long rows = (long long)n * ndof;
    PXG_CHECK(rowptr[0] == 0, PXG_ERR_ARG, "pxg_set_csr: rowptr[0] must be 0");
    for (long long r = 0; r < rows; r++)
        PXG_CHECK(rowptr[r + 1] >= rowptr[r], PXG_ERR_ARG, "pxg_set_csr: row offsets must not decrease");
    const long long nnz = rowptr[rows];
    PXG_CHECK(nnz > 0, PXG_ERR_ARG, "pxg_set_csr: empty pattern");
    std::vector<int> off((size_t)n + 1), nbr((size_t)ctx->nb);
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PXG_HIP(hipMemcpyAsync(off.data(), ctx->off.p, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(nbr.data(), ctx->nbr.p, sizeof(int) * (size_t)ctx->nb, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    // Every node owns ndof rows with one column pattern: whole ndof x ndof
    // blocks, sorted, holding the diagonal block and the family's blocks
    // (SparseMatrix.cpp:90-115).  At most one more block (the ghost's mirror)
    // may be present; it receives no term and is zeroed by the assembly.
    std::vector<int> slot((size_t)ctx->nb), diag((size_t)n), mslot((size_t)n, -1);
    int maxblk = 0;
    std::vector<int> blk;
    for (int i = 0; i < n; i++) {
        const long long r0 = rowptr[(size_t)i * ndof], len = rowptr[(size_t)i * ndof + 1] - r0;
        PXG_CHECK(len % ndof == 0, PXG_ERR_ARG, "pxg_set_csr: a row does not hold whole dof blocks");
        const int nbk = (int)(len / ndof);
        blk.resize((size_t)nbk);
        for (int k = 0; k < nbk; k++) {
            const int c = cols[r0 + (long long)k * ndof];
            PXG_CHECK(c >= 0 && c < rows && c % ndof == 0 && (k == 0 || c / ndof > blk[(size_t)k - 1]), PXG_ERR_ARG,
                      "pxg_set_csr: columns must be sorted whole dof blocks");
            blk[(size_t)k] = c / ndof;
            for (int b = 1; b < ndof; b++)
                PXG_CHECK(cols[r0 + (long long)k * ndof + b] == c + b, PXG_ERR_ARG,
                          "pxg_set_csr: columns must be sorted whole dof blocks");
        }
        for (int a = 1; a < ndof; a++) {
            const long long ra = rowptr[(size_t)i * ndof + a];
            PXG_CHECK(rowptr[(size_t)i * ndof + a + 1] - ra == len, PXG_ERR_ARG,
                      "pxg_set_csr: the rows of one node must share their pattern");
            for (long long q = 0; q < len; q++)
                PXG_CHECK(cols[ra + q] == cols[r0 + q], PXG_ERR_ARG,
                          "pxg_set_csr: the rows of one node must share their pattern");
        }
        auto pos = [&](int node0) {
            const auto it = std::lower_bound(blk.begin(), blk.end(), node0);
            return (it != blk.end() && *it == node0) ? (int)(it - blk.begin()) : -1;
        };
        const int pd = pos(i);
        PXG_CHECK(pd >= 0, PXG_ERR_ARG, "pxg_set_csr: a row lacks its diagonal block");
        diag[(size_t)i] = (int)(r0 + (long long)pd * ndof);
        std::vector<char> used((size_t)nbk, 0);
        used[(size_t)pd] = 1;
        for (int t = off[(size_t)i]; t < off[(size_t)i + 1]; t++) {
            const int pj = pos(nbr[(size_t)t] - 1);
            PXG_CHECK(pj >= 0, PXG_ERR_ARG, "pxg_set_csr: a family member is missing from its row");
            slot[(size_t)t] = (int)(r0 + (long long)pj * ndof);
            used[(size_t)pj] = 1;
        }
        for (int k = 0; k < nbk; k++) {
            if (used[(size_t)k]) continue;
            PXG_CHECK(mslot[(size_t)i] < 0, PXG_ERR_ARG,
                      "pxg_set_csr: a row holds more than one block outside the node's family");
            mslot[(size_t)i] = (int)(r0 + (long long)k * ndof);
        }
        maxblk = std::max(maxblk, nbk);
    }
    PXG_FAULT(ctx, "pxg_set_csr");
    ctx->csrbc.set = 0;
    ctx->csr_gen++;
    ctx->solve.own = 0;                 // the solver follows the new implicit CSR
    PXG_HIP(ctx->rowptr.alloc((size_t)rows + 1));
    PXG_HIP(ctx->cols.alloc((size_t)nnz));
    PXG_HIP(ctx->slot.alloc((size_t)ctx->nb));
    PXG_HIP(ctx->diag.alloc((size_t)n));
    PXG_HIP(ctx->mslot.alloc((size_t)n));
    PXG_HIP(hipMemcpyAsync(ctx->rowptr.p, rowptr, sizeof(int) * ((size_t)rows + 1), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->cols.p, cols, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->slot.p, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->diag.p, diag.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->mslot.p, mslot.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    ctx->nnz = nnz;
    ctx->csr_ndof = ndof;
    ctx->csr_maxblk = maxblk;
    ctx->ch_ready = 0;
    return PXG_OK;
}

extern "C" int pxg_get_families(pxg_ctx *ctx, int *off, int *nbr)
{
    PXG_CHECK(ctx && ctx->nb > 0, PXG_ERR_STATE, "pxg_get_families: no families");
    PXG_HIP(hipSetDevice(ctx->device));
    if (off) PXG_HIP(hipMemcpyAsync(off, ctx->off.p, sizeof(int) * ((size_t)ctx->n + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (nbr) PXG_HIP(hipMemcpyAsync(nbr, ctx->nbr.p, sizeof(int) * (size_t)ctx->nb, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_set_cracks(pxg_ctx *ctx, int nseg, const double *seg, int nquad, const double *quad)
{
    PXG_CHECK(ctx && nseg >= 0 && nquad >= 0 && (nseg == 0 || seg) && (nquad == 0 || quad),
              PXG_ERR_ARG, "pxg_set_cracks: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    ctx->nseg = nseg; ctx->nquad = nquad;
    PXG_HIP(ctx->crack_seg.alloc((size_t)4 * (nseg > 0 ? nseg : 1)));
    PXG_HIP(ctx->crack_quad.alloc((size_t)12 * (nquad > 0 ? nquad : 1)));
    if (nseg) PXG_HIP(hipMemcpyAsync(ctx->crack_seg.p, seg, sizeof(double) * 4 * nseg, hipMemcpyHostToDevice, ctx->stream));
    if (nquad) PXG_HIP(hipMemcpyAsync(ctx->crack_quad.p, quad, sizeof(double) * 12 * nquad, hipMemcpyHostToDevice, ctx->stream));
    if (ctx->nb > 0) {
        if (nseg + nquad > 0)
            hipLaunchKernelGGL(k_seed, dim3(grid_of(ctx->n, kConvNodes)), dim3(kConvBlock), 0, ctx->stream, ctx->coords.p,
                               ctx->mirror.p, ctx->n, ctx->off.p, ctx->nbr.p, ctx->crack_seg.p, nseg,
                               ctx->crack_quad.p, nquad, ctx->seed.p);
        else
            PXG_HIP(hipMemsetAsync(ctx->seed.p, 1, (size_t)ctx->nb, ctx->stream));
        PXG_HIP(hipGetLastError());
    }
    invalidate(ctx, INV_CRACKS);
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_get_seed(pxg_ctx *ctx, unsigned char *seed)
{
    PXG_CHECK(ctx && ctx->nb > 0 && seed, PXG_ERR_STATE, "pxg_get_seed: no families");
    PXG_HIP(hipSetDevice(ctx->device));
    PXG_HIP(hipMemcpyAsync(seed, ctx->seed.p, (size_t)ctx->nb, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_build_csr(pxg_ctx *ctx, int ndof, long long *nnz_out)
{
    PXG_CHECK(ctx && ctx->nb > 0, PXG_ERR_STATE, "pxg_build_csr: no families");
    PXG_CHECK(ndof >= 1 && ndof <= 8, PXG_ERR_ARG, "pxg_build_csr: ndof must be 1..8");
    PXG_FAULT(ctx, "pxg_build_csr");
    PXG_HIP(hipSetDevice(ctx->device));
    ctx->csrbc.set = 0;                 // a BC row program refers to the old pattern
    ctx->csr_gen++;
    ctx->solve.own = 0;                 // the solver follows the new implicit CSR
    hipStream_t st = ctx->stream;
    const int n = ctx->n;
    const long long rows = (long long)n * ndof;
    dbuf<int> rc, flag;
    PXG_HIP(rc.alloc(rows));
    PXG_HIP(flag.alloc(2));
    PXG_HIP(hipMemsetAsync(flag.p, 0, 2 * sizeof(int), st));
    hipLaunchKernelGGL(k_csr_count, dim3(grid_of(n, 64)), dim3(64), 0, st, ctx->off.p, ctx->nbr.p,
                       ctx->mirror.p, n, ndof, rc.p, flag.p);
    int fl[2] = {0, 0};
    PXG_HIP(hipMemcpyAsync(fl, flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    PXG_CHECK(!fl[0], PXG_ERR_ARG, "pxg_build_csr: a family exceeds the 510-bond row limit");
    ctx->csr_maxblk = fl[1];
    PXG_HIP(ctx->rowptr.alloc(rows + 1));
    long long nnz = 0;
    int r = scan_counts(st, rc.p, rows, ctx->rowptr.p, &nnz);
    if (r) return r;
    ctx->nnz = nnz;
    ctx->csr_ndof = ndof;
    PXG_HIP(ctx->cols.alloc(nnz));
    PXG_HIP(ctx->slot.alloc(ctx->nb));
    PXG_HIP(ctx->diag.alloc(n));
    PXG_HIP(ctx->mslot.alloc(n));
    hipLaunchKernelGGL(k_csr_fill, dim3(grid_of(n, 64)), dim3(64), 0, st, ctx->off.p, ctx->nbr.p,
                       ctx->mirror.p, n, ndof, ctx->rowptr.p, ctx->cols.p, ctx->slot.p, ctx->diag.p,
                       ctx->mslot.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(st));
    ctx->ch_ready = 0;
    if (nnz_out) *nnz_out = nnz;
    return PXG_OK;
}

extern "C" int pxg_get_csr(pxg_ctx *ctx, int *rowptr, int *cols)
{
    PXG_CHECK(ctx && ctx->nnz > 0, PXG_ERR_STATE, "pxg_get_csr: no CSR pattern");
    PXG_HIP(hipSetDevice(ctx->device));
    const size_t rows = (size_t)ctx->n * ctx->csr_ndof;
    if (rowptr) PXG_HIP(hipMemcpyAsync(rowptr, ctx->rowptr.p, sizeof(int) * (rows + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (cols) PXG_HIP(hipMemcpyAsync(cols, ctx->cols.p, sizeof(int) * (size_t)ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}
