// pxg_api.hip -- context lifetime, error reporting and fault injection.
#include "pxg_internal.h"

namespace pxg {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

void invalidate(pxg_ctx *ctx, int what)
{
    if (what & INV_OPS) {
        // bond data sized and computed for the old cloud / families / operators
        ctx->ops_gen++;
        ctx->lap_kind = LAP_NONE;
        ctx->ch_ready = 0;
        ctx->frac.ready = 0;
        ctx->dyn.ready = 0;
        ctx->sch.ready = 0;
        ctx->fsch.ready = 0;
        ctx->mob_set = 0;
    }
    if (what & INV_CRACKS) {
        // element health started from the old crack seed
        ctx->frac.ready = 0;
        ctx->dyn.ready = 0;
        ctx->fsch.ready = 0;
    }
}
} // namespace pxg

using namespace pxg;

extern "C" int pxg_create(pxg_ctx **out, int device)
{
    PXG_CHECK(out, PXG_ERR_ARG, "pxg_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    PXG_HIP(hipGetDeviceCount(&ndev));
    PXG_CHECK(device >= 0 && device < ndev, PXG_ERR_ARG, "pxg_create: no such device");
    PXG_HIP(hipSetDevice(device));
    {
        const int rc = scratch_init(device);
        if (rc) return rc;
    }
    pxg_ctx *c = new pxg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("pxg_create: ") + hipGetErrorString(e));
        return PXG_ERR_HIP;
    }
    *out = c;
    return PXG_OK;
}

template <typename T> static void rel(dbuf<T> &b) { b.release(); }

extern "C" void pxg_destroy(pxg_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    rel(c->coords); rel(c->xp); rel(c->vols); rel(c->mirror); rel(c->off); rel(c->nbr);
    rel(c->crack_seg); rel(c->crack_quad); rel(c->seed); rel(c->ops); rel(c->flags);
    rel(c->rowptr); rel(c->cols); rel(c->slot); rel(c->diag); rel(c->mslot); rel(c->lap);
    rel(c->rev); rel(c->isghost); rel(c->su); rel(c->suo); rel(c->sk);
    rel(c->mob); rel(c->xrev); rel(c->xlap); rel(c->xmob); rel(c->xvel); rel(c->xghost); rel(c->xhealth);
    rel(c->sr); rel(c->sf);
    rel(c->csrbc.row); rel(c->csrbc.dslot); rel(c->csrbc.pslot); rel(c->csrbc.pdof);
    rel(c->csrbc.kind); rel(c->csrbc.val0); rel(c->csrbc.vel);
    {
        pxg::SolveState &S = c->solve;
        S.drop_graphs();
        rel(S.F); rel(S.Li); rel(S.Ui); rel(S.y); rel(S.z); rel(S.w); rel(S.inv); rel(S.fb); rel(S.al);
        rel(S.zpiv); rel(S.loc); rel(S.lp); rel(S.part); rel(S.norms);
    }
    rel(c->dyn.health); rel(c->dyn.vel); rel(c->dyn.uprev);
    rel(c->fsch.health); rel(c->fsch.lapc); rel(c->fsch.vel); rel(c->fsch.uprev);
    FracState &F = c->frac;
    rel(F.P64); rel(F.P32); rel(F.health); rel(F.break_step); rel(F.u[0]); rel(F.u[1]);
    rel(F.perm); rel(F.rcnt); rel(F.ulist); rel(F.nrec); rel(F.brec); rel(F.gslot); rel(F.bovf); rel(F.ghist64); rel(F.bc.prog_of); rel(F.bsh); rel(F.pofs); rel(F.health_csr); rel(F.bs_csr);
    rel(F.preset_mask); rel(F.bc.node); rel(F.bc.src); rel(F.bc.mode); rel(F.bc.val0);
    rel(F.bc.vel); rel(F.gmax); rel(F.ghist); rel(F.gstate); rel(F.nodeE); rel(F.Phat);
    rel(F.rev); rel(F.accel); rel(F.anybroke); rel(F.bc.inv_off); rel(F.bc.inv_q);
    if (F.gstate_host) (void)hipHostFree(F.gstate_host);
    for (hipEvent_t e : F.ev) (void)hipEventDestroy(e);
    F.ev.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *pxg_last_error(void) { return g_err.c_str(); }

extern "C" int pxg_set_fault(pxg_ctx *ctx, int code)
{
    PXG_CHECK(ctx, PXG_ERR_ARG, "pxg_set_fault: ctx is NULL");
    ctx->fault = code;
    return PXG_OK;
}

extern "C" int pxg_sync(pxg_ctx *ctx)
{
    PXG_CHECK(ctx, PXG_ERR_ARG, "pxg_sync: ctx is NULL");
    PXG_HIP(hipSetDevice(ctx->device));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}
