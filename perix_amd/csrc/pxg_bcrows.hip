// pxg_bcrows.hip -- boundary-condition row elimination of the implicit CSR
// system on the device (BCSystem::applyBCs, BCSystem.cpp:212-233).
//
// The reference walks its BC list in order and, per constrained dof row g,
//   scale = bcRowScale(K, g)                       (BCBase.h:33-39)
//   zero row g, K(g,g) = scale                     (SparseMatrix.cpp:333-350)
//   Dirichlet, mirror bulk b, not direct:  K(g,b) = scale,
//       R(g) = scale (2 g(t) - U(g) - U(b))        (DirichletBC.cpp:208-224)
//   Dirichlet direct or no mirror:  R(g) = scale (g(t) - U(g))
//   Neumann:  K(g,b) = -scale, R(g) = scale (U(b) - U(g))
//                                                  (NeumannBC.cpp:127-151)
// with g(t) = value + velocity t and the Dirichlet box filter
// (DirichletBC.cpp:89-98).  Rows are independent: an entry reads and writes
// row g of K and R(g) only.  When several entries reach one row, the first
// leaves |K(g,g)| >= 1e-100 (or NaN), so every later one takes the same
// scale, and each later one zeroes the row and rewrites R(g): the result is
// the last entry's formula with the scale of the row as assembled.  The set-up
// therefore resolves the list into one program entry per row (the last entry
// that reaches it) and the kernel applies all rows at once, one wavefront per
// row.
#include <algorithm>
#include <vector>
#include "pxg_internal.h"

using namespace pxg;

namespace {

constexpr int kRowsPerBlock = 4;   // one 64-lane wave per constrained row

__global__ __launch_bounds__(256) void k_bc_rows(int count, const int *__restrict__ prow,
                                                 const int *__restrict__ pdslot,
                                                 const int *__restrict__ ppslot,
                                                 const int *__restrict__ ppdof,
                                                 const signed char *__restrict__ pkind,
                                                 const double *__restrict__ pval0,
                                                 const double *__restrict__ pvel, double t,
                                                 const int *__restrict__ rowptr,
                                                 const double *__restrict__ U, double *K, double *R)
{
    const int q = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (q >= count) return;             // uniform per wave
    const int g = prow[q], d = pdslot[q], p = ppslot[q];
    const int s = rowptr[g], e = rowptr[g + 1];
    // largest |K(g,k)|; NaN entries never win, as in getRowMaxAbs
    double m = 0.0;
    for (int k = s + lane; k < e; k += 64) {
        const double a = fabs(K[k]);
        if (a > m) m = a;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double w = __shfl_xor(m, o, 64);
        m = (w > m) ? w : m;
    }
    double scale = K[d];
    if (fabs(scale) < 1.0e-100) scale = m;
    if (fabs(scale) < 1.0e-100) scale = 1.0;
    // every lane has read the row (the reduction above joins them); the
    // diagonal and the partner slot are written once, by lane 0
    for (int k = s + lane; k < e; k += 64)
        if (k != d && k != p) K[k] = 0.0;
    if (lane == 0) {
        K[d] = scale;
        const double ug = U[g];
        const int kind = pkind[q];
        if (kind == BCROW_MIRROR) {
            K[p] = -scale;
            R[g] = scale * (U[ppdof[q]] - ug);
        } else {
            const double value = pval0[q] + pvel[q] * t;
            if (kind == BCROW_REFLECT) {
                K[p] = scale;
                R[g] = scale * (2.0 * value - ug - U[ppdof[q]]);
            } else {
                R[g] = scale * (value - ug);
            }
        }
    }
}

// CSR slot of (row, col) in the sorted row, -1 if not stored
// (SparseMatrix::findEntry, SparseMatrix.cpp:187-214)
int find_slot(const std::vector<int> &rp, const std::vector<int> &cols, int row, int col)
{
    const auto b = cols.begin() + rp[(size_t)row], e = cols.begin() + rp[(size_t)row + 1];
    const auto it = std::lower_bound(b, e, col);
    return (it != e && *it == col) ? (int)(it - cols.begin()) : -1;
}

} // namespace

extern "C" int pxg_csr_set_bcs(pxg_ctx *ctx, int nbc, const pxg_bc *bcs)
{
    PXG_CHECK(ctx && ctx->nnz > 0 && ctx->csr_ndof > 0, PXG_ERR_STATE,
              "pxg_csr_set_bcs: call pxg_build_csr first");
    PXG_CHECK(nbc >= 0 && (nbc == 0 || bcs), PXG_ERR_ARG, "pxg_csr_set_bcs: bad arguments");
    PXG_FAULT(ctx, "pxg_csr_set_bcs");
    PXG_HIP(hipSetDevice(ctx->device));
    CsrBcProgram &P = ctx->csrbc;
    P.set = 0;
    const int n = ctx->n, ndof = ctx->csr_ndof;
    const size_t rows = (size_t)n * ndof;
    std::vector<int> rp(rows + 1), cols((size_t)ctx->nnz);
    hipStream_t st = ctx->stream;
    PXG_HIP(hipMemcpyAsync(rp.data(), ctx->rowptr.p, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(cols.data(), ctx->cols.p, sizeof(int) * cols.size(), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));

    struct Entry { int row, dslot, pslot, pdof; signed char kind; double val0, vel; };
    std::vector<int> at(rows, -1);
    std::vector<Entry> ent;
    for (int e = 0; e < nbc; e++) {
        const pxg_bc &b = bcs[e];
        PXG_CHECK((b.type == 0 || b.type == 1) && b.ncomp >= 0 && b.ncomp <= 3 &&
                      (b.nnodes == 0 || b.nodes),
                  PXG_ERR_ARG, "pxg_csr_set_bcs: bad BC entry");
        for (int q = 0; q < b.nnodes; q++) {
            const int id = b.nodes[q];
            PXG_CHECK(id >= 1 && id <= n, PXG_ERR_ARG, "pxg_csr_set_bcs: node outside the PD mesh");
            const int mb = ctx->hmirror[(size_t)id - 1];
            if (b.type == 0) {
                PXG_CHECK(mb >= 1 && mb <= n, PXG_ERR_ARG,
                          "pxg_csr_set_bcs: neumann node has no mirror bulk node (NeumannBC.cpp:67-85)");
            } else if (b.has_box) {
                bool in = true;
                for (int ax = 0; ax < 3 && in; ax++) {
                    const double lo = b.box[2 * ax], hi = b.box[2 * ax + 1];
                    if (lo > hi) continue;
                    const double x = ctx->hcoords[3 * ((size_t)id - 1) + ax];
                    in = !(x < lo || x > hi);
                }
                if (!in) continue;
            }
            for (int k = 0; k < b.ncomp; k++) {
                const int c = b.comp[k];
                PXG_CHECK(c >= 0 && c < ndof, PXG_ERR_ARG, "pxg_csr_set_bcs: component out of range");
                const int g = (id - 1) * ndof + c;
                Entry x{g, find_slot(rp, cols, g, g), -1, -1, (signed char)BCROW_DIRECT, 0.0, 0.0};
                PXG_CHECK(x.dslot >= 0, PXG_ERR_ARG,
                          "pxg_csr_set_bcs: constrained row has no stored diagonal (SparseMatrix.cpp:341-346)");
                if (b.type == 0) {
                    x.kind = BCROW_MIRROR;
                } else {
                    x.val0 = b.value[k];
                    x.vel = b.velocity[k];
                    if (mb >= 1 && !b.direct) x.kind = BCROW_REFLECT;
                }
                if (x.kind != BCROW_DIRECT) {
                    x.pdof = (mb - 1) * ndof + c;
                    x.pslot = find_slot(rp, cols, g, x.pdof);
                    PXG_CHECK(x.pslot >= 0, PXG_ERR_ARG,
                              "pxg_csr_set_bcs: mirror entry not in the CSR pattern (SparseMatrix.cpp:225-232)");
                }
                if (at[(size_t)g] < 0) {
                    at[(size_t)g] = (int)ent.size();
                    ent.push_back(x);
                } else {
                    ent[(size_t)at[(size_t)g]] = x;     // a later entry rewrites the row
                }
            }
        }
    }
    std::sort(ent.begin(), ent.end(), [](const Entry &a, const Entry &b) { return a.row < b.row; });
    const size_t m = ent.size();
    std::vector<int> row(m), dslot(m), pslot(m), pdof(m);
    std::vector<signed char> kind(m);
    std::vector<double> val0(m), vel(m);
    for (size_t q = 0; q < m; q++) {
        row[q] = ent[q].row; dslot[q] = ent[q].dslot; pslot[q] = ent[q].pslot; pdof[q] = ent[q].pdof;
        kind[q] = ent[q].kind; val0[q] = ent[q].val0; vel[q] = ent[q].vel;
    }
    P.count = (int)m;
    P.chained = 0;
    for (size_t q = 0; q < m; q++)
        if (pdof[q] >= 0 && at[(size_t)pdof[q]] >= 0) P.chained = 1;
    if (m > 0) {
        PXG_HIP(P.row.alloc(m)); PXG_HIP(P.dslot.alloc(m)); PXG_HIP(P.pslot.alloc(m));
        PXG_HIP(P.pdof.alloc(m)); PXG_HIP(P.kind.alloc(m)); PXG_HIP(P.val0.alloc(m));
        PXG_HIP(P.vel.alloc(m));
        PXG_HIP(hipMemcpyAsync(P.row.p, row.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.dslot.p, dslot.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.pslot.p, pslot.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.pdof.p, pdof.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.kind.p, kind.data(), m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.val0.p, val0.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(P.vel.p, vel.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
        PXG_HIP(hipStreamSynchronize(st));
    }
    P.set = 1;
    return PXG_OK;
}

extern "C" int pxg_csr_bc_rows(pxg_ctx *ctx, int *count)
{
    PXG_CHECK(ctx && ctx->csrbc.set, PXG_ERR_STATE, "pxg_csr_bc_rows: call pxg_csr_set_bcs first");
    PXG_CHECK(count, PXG_ERR_ARG, "pxg_csr_bc_rows: bad arguments");
    *count = ctx->csrbc.count;
    return PXG_OK;
}

extern "C" int pxg_csr_apply_bcs_dev(pxg_ctx *ctx, double t, const double *u_dev, double *kvals_dev,
                                     double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->csrbc.set, PXG_ERR_STATE, "pxg_csr_apply_bcs: call pxg_csr_set_bcs first");
    PXG_CHECK(u_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_csr_apply_bcs: bad arguments");
    PXG_FAULT(ctx, "pxg_csr_apply_bcs");
    PXG_HIP(hipSetDevice(ctx->device));
    const CsrBcProgram &P = ctx->csrbc;
    if (P.count == 0) return PXG_OK;
    const unsigned grid = (unsigned)((P.count + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL(k_bc_rows, dim3(grid), dim3(256), 0, ctx->stream, P.count, P.row.p, P.dslot.p,
                       P.pslot.p, P.pdof.p, P.kind.p, P.val0.p, P.vel.p, t, ctx->rowptr.p, u_dev,
                       kvals_dev, rhs_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_csr_apply_bcs(pxg_ctx *ctx, double t, const double *u, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->csrbc.set, PXG_ERR_STATE, "pxg_csr_apply_bcs: call pxg_csr_set_bcs first");
    PXG_CHECK(u && kvals && rhs, PXG_ERR_ARG, "pxg_csr_apply_bcs: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t rows = (size_t)ctx->n * ctx->csr_ndof, nnz = (size_t)ctx->nnz;
    PXG_HIP(ctx->su.alloc(rows));
    PXG_HIP(ctx->sk.alloc(nnz));
    PXG_HIP(ctx->sr.alloc(rows));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * rows, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->sk.p, kvals, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->sr.p, rhs, sizeof(double) * rows, hipMemcpyHostToDevice, st));
    int rc = pxg_csr_apply_bcs_dev(ctx, t, ctx->su.p, ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}
