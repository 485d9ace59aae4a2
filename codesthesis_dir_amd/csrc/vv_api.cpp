// vv_api.cpp -- C ABI of libvvmi (declared in include/vvmi.h): run
// lifetime, host/device field transfer, the reference routines as device
// operations, and the fused multi-step driver.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vv_internal.h"

namespace vv {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return VV_EHIP;
}

// rb.f:3046-3064 elimination factors (same operations as the reference)
TdmaTab make_tdma(const std::vector<double> &a, const std::vector<double> &b,
                  const std::vector<double> &c)
{
    TdmaTab t;
    const int n = (int)b.size();
    t.n = n;
    t.fact.assign(n, 0.0);
    t.bp = b;
    t.c = c;
    for (int j = 1; j < n; ++j) {
        const double fact = a[j] / t.bp[j - 1];
        t.fact[j] = fact;
        t.bp[j] = t.bp[j] - fact * c[j - 1];
    }
    return t;
}

// rb.f:3066-3116 elimination factors; arrays are 1-based (index 0 unused)
PtdmaTab make_ptdma(const std::vector<double> &a0, const std::vector<double> &b0,
                    const std::vector<double> &c0)
{
    PtdmaTab t;
    const int n = (int)b0.size();
    t.n = n;
    std::vector<double> a(n + 1), b(n + 1), c(n + 1);
    for (int j = 1; j <= n; ++j) { a[j] = a0[j - 1]; b[j] = b0[j - 1]; c[j] = c0[j - 1]; }
    t.f1.assign(n + 1, 0.0);
    t.f2.assign(n + 1, 0.0);
    double fact;
    for (int j = 2; j <= n - 2; ++j) {
        fact = a[j] / b[j - 1];
        t.f1[j] = fact;
        b[j] = b[j] - fact * c[j - 1];
        a[j] = -fact * a[j - 1];
        fact = c[n] / b[j - 1];
        t.f2[j] = fact;
        c[n] = -fact * c[j - 1];
        b[n] = b[n] - fact * a[j - 1];
    }
    fact = a[n - 1] / b[n - 2];
    t.fA = fact;
    b[n - 1] = b[n - 1] - fact * c[n - 2];
    c[n - 1] = c[n - 1] - fact * a[n - 2];
    fact = c[n] / b[n - 2];
    t.fB = fact;
    a[n] = a[n] - fact * c[n - 2];
    b[n] = b[n] - fact * a[n - 2];
    fact = a[n] / b[n - 1];
    t.fC = fact;
    b[n] = b[n] - fact * c[n - 1];
    t.a = a;
    t.b = b;
    t.c = c;
    return t;
}

int prof_begin(vv_run *r, hipEvent_t *a)
{
    *a = nullptr;
    if (!r->prof_on) return VV_OK;
    if (r->ev_used == r->ev_pool.size()) {
        hipEvent_t e;
        VV_HIP(hipEventCreate(&e));
        r->ev_pool.push_back(e);
    }
    *a = r->ev_pool[r->ev_used++];
    VV_HIP(hipEventRecord(*a, r->stream));
    return VV_OK;
}

int prof_end(vv_run *r, int kid, hipEvent_t a, double bytes)
{
    if (!r->prof_on || !a) return VV_OK;
    if (r->ev_used == r->ev_pool.size()) {
        hipEvent_t e;
        VV_HIP(hipEventCreate(&e));
        r->ev_pool.push_back(e);
    }
    hipEvent_t b = r->ev_pool[r->ev_used++];
    VV_HIP(hipEventRecord(b, r->stream));
    r->prof_pending.push_back({kid, a, b, bytes});
    return VV_OK;
}

// fold the recorded event pairs into the per-kernel totals (stream idle)
static int prof_resolve(vv_run *r)
{
    for (const auto &p : r->prof_pending) {
        float ms = 0.f;
        VV_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        r->prof_launch[p.kid] += 1;
        r->prof_ms[p.kid] += ms;
        r->prof_bytes[p.kid] += p.bytes;
    }
    r->prof_pending.clear();
    r->ev_used = 0;
    return VV_OK;
}

}  // namespace vv

using namespace vv;

namespace {

using Pending = vv_run::Pend;

// pending Robert filter: between the explicit step and update the filtered
// current level lives in Cf/thf
Pending &pending(vv_run *r) { return r->pend; }

int alloc_field(vv_run *r, double **p)
{
    if (hipMalloc(p, r->fbytes) != hipSuccess) {
        set_error("device allocation failed");
        return VV_ENOMEM;
    }
    VV_HIP(hipMemsetAsync(*p, 0, r->fbytes, r->stream));
    return VV_OK;
}

double *field_ptr(vv_run *r, int f)
{
    Pending &pd = pending(r);
    switch (f) {
    case VV_C: return pd.c ? r->Cf : r->C;
    case VV_CP: return r->Cp;
    case VV_CN: return r->Cn;
    case VV_TH: return pd.t ? r->thf : r->th;
    case VV_THP: return r->thp;
    case VV_THN: return r->thn;
    case VV_UEL: return r->uel;
    case VV_VEL: return r->vel;
    case VV_UELP: return r->uelp;
    case VV_VELP: return r->velp;
    case VV_VRH: return r->vrh;
    case VV_URH: return r->urh;
    case VV_MEAN_T2: return r->mT2;
    case VV_MEAN_T: return r->mT;
    case VV_MEAN_U2: return r->mU2;
    case VV_MEAN_V2: return r->mV2;
    default: return nullptr;
    }
}

// the device recurrences divide by the pivots with a quotient form that
// needs positive divisors
bool positive(const std::vector<double> &v, size_t from)
{
    for (size_t j = from; j < v.size(); ++j)
        if (!(v[j] > 0.0)) return false;
    return true;
}

int upload_tab(vv_run *r, double **dst, const std::vector<double> &h)
{
    VV_HIP(hipMalloc(dst, h.size() * sizeof(double)));
    // synchronous: the caller's vector is freed on return
    VV_HIP(hipMemcpy(*dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    (void)r;
    return VV_OK;
}

int build_tables(vv_run *r)
{
    const int nr = r->nr, np = r->np;
    {   // Cr, thr at vorticity points: 1-4-1, n = nr-1 (rb.f:810-826)
        const int n = nr - 1;
        std::vector<double> a(n, 1.0), b(n, 4.0), c(n, 1.0);
        TdmaTab t = make_tdma(a, b, c);
        if (!positive(t.bp, 0)) { set_error("vv_create: non-positive tdma pivot"); return VV_EINVAL; }
        std::vector<double> h;
        h.insert(h.end(), t.fact.begin(), t.fact.end());
        h.insert(h.end(), t.bp.begin(), t.bp.end());
        h.insert(h.end(), t.c.begin(), t.c.end());
        for (double v : t.bp) h.push_back(1.0 / v);
        int st = upload_tab(r, &r->tab_r1, h);
        if (st) return st;
    }
    {   // thr at temperature points with closure rows (rb.f:1244-1262)
        const int n = nr;
        std::vector<double> a(n, 1.0), b(n, 4.0), c(n, 1.0);
        b[0] = 1.0; c[0] = 2.0;
        a[n - 1] = 2.0; b[n - 1] = 1.0;
        TdmaTab t = make_tdma(a, b, c);
        if (!positive(t.bp, 0)) { set_error("vv_create: non-positive tdma pivot"); return VV_EINVAL; }
        std::vector<double> h;
        h.insert(h.end(), t.fact.begin(), t.fact.end());
        h.insert(h.end(), t.bp.begin(), t.bp.end());
        h.insert(h.end(), t.c.begin(), t.c.end());
        for (double v : t.bp) h.push_back(1.0 / v);
        int st = upload_tab(r, &r->tab_rT, h);
        if (st) return st;
    }
    {   // cyclic 1-4-1 on np+1 points (rb.f:833-845)
        const int n = np + 1;
        std::vector<double> a(n, 1.0), b(n, 4.0), c(n, 1.0);
        PtdmaTab t = make_ptdma(a, b, c);
        if (!positive(t.b, 1)) { set_error("vv_create: non-positive ptdma pivot"); return VV_EINVAL; }
        std::vector<double> h;
        h.insert(h.end(), t.f1.begin(), t.f1.end());
        h.insert(h.end(), t.f2.begin(), t.f2.end());
        h.insert(h.end(), t.a.begin(), t.a.end());
        h.insert(h.end(), t.b.begin(), t.b.end());
        h.insert(h.end(), t.c.begin(), t.c.end());
        h.push_back(t.fA);
        h.push_back(t.fB);
        h.push_back(t.fC);
        for (double v : t.b) h.push_back(v != 0.0 ? 1.0 / v : 0.0);   // entry 0 is unused
        int st = upload_tab(r, &r->tab_p, h);
        if (st) return st;
    }
    return VV_OK;
}

void rotate(vv_run *r)
{
    double *oC = r->C, *oCp = r->Cp;
    r->C = r->Cn;
    r->Cp = r->Cf;
    r->Cn = oCp;
    r->Cf = oC;
    double *oT = r->th, *oTp = r->thp;
    r->th = r->thn;
    r->thp = r->thf;
    r->thn = oTp;
    r->thf = oT;
}

}  // namespace

extern "C" {

const char *vv_last_error(void) { return g_err.c_str(); }

int vv_device_count(int *n)
{
    if (!n) return VV_EINVAL;
    VV_HIP(hipGetDeviceCount(n));
    return VV_OK;
}

int vv_create(const vv_config *cfg, int device, vv_run **out, int *nlev)
{
    if (!cfg || !out) { set_error("null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->nr < 8 || cfg->np < 8) {
        set_error("need nr >= 8 and np >= 8");
        return VV_EINVAL;
    }
    // a workgroup of the compact scheme holds one whole line in LDS, and the
    // line solves of the Poisson levels hold lines of nr rows
    if (cfg->np > 16383 || cfg->nr > 2048) {
        set_error("need np <= 16383 and nr <= 2048 (lines are held in LDS)");
        return VV_EINVAL;
    }
    if (!(cfg->dt > 0) || !(cfg->rmax > cfg->r0) || !(cfg->Pr > 0) ||
        (cfg->geom == VV_GEOM_BOX && !(cfg->width > 0)) ||
        (cfg->geom == VV_GEOM_ANNULUS && (!(cfg->r0 > 0) || cfg->Re == 0.0)) ||
        (cfg->geom != VV_GEOM_BOX && cfg->geom != VV_GEOM_ANNULUS)) {
        set_error("bad physical parameters");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vv_run *r = new vv_run();
    r->cfg = *cfg;
    if (r->cfg.rtol <= 0) r->cfg.rtol = 1e-10;
    if (r->cfg.max_cycles <= 0) r->cfg.max_cycles = 100;
    if (r->cfg.nu1 <= 0) r->cfg.nu1 = 1;
    if (r->cfg.nu2 <= 0) r->cfg.nu2 = 1;
    if ((r->cfg.nu1 + r->cfg.nu2) % 2 != 0) {
        // the sweeps alternate between two buffers per level; an even count
        // returns the solution to the first one at the end of each cycle
        set_error("nu1 + nu2 must be even");
        delete r;
        return VV_EINVAL;
    }
    if (r->cfg.robert < 0) r->cfg.robert = 0.02;
    r->device = device;
    r->nr = cfg->nr;
    r->np = cfg->np;
    r->ld = ((cfg->nr + 2 + 7) / 8) * 8;
    r->ncols = cfg->np + 1;
    r->dr = (cfg->rmax - cfg->r0) / cfg->nr;
    if (cfg->geom == VV_GEOM_ANNULUS) {
        const double pi = std::acos(-1.0);           // ann.f:121-124
        r->dp = 2.0 * pi / (double)cfg->np;
    } else {
        r->dp = cfg->width / cfg->np;
    }
    r->fbytes = (size_t)r->ld * r->ncols * sizeof(double);
    int st = VV_OK;
    auto fail = [&](int s) { vv_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    if (hipEventCreate(&r->ev0) != hipSuccess || hipEventCreate(&r->ev1) != hipSuccess) return fail(VV_EHIP);
    double **fields[] = {&r->C, &r->Cp, &r->Cn, &r->Cf, &r->th, &r->thp, &r->thn, &r->thf,
                         &r->uel, &r->vel, &r->uelp, &r->velp, &r->vrh, &r->urh,
                         &r->dCr, &r->dCh, &r->dthrC, &r->dthhC, &r->dthrT, &r->dthhT,
                         &r->mT2, &r->mT, &r->mU2, &r->mV2};
    for (double **f : fields)
        if ((st = alloc_field(r, f))) return fail(st);
    if ((st = build_tables(r))) return fail(st);
    {   // row radii exactly as the reference forms them (ann.f:918, 1331)
        std::vector<double> rc(r->ld, 1.0), rt(r->ld, 1.0);
        for (int i = 0; i <= r->nr + 1; ++i) {
            rc[i] = cfg->r0 + (double)i * r->dr;
            rt[i] = cfg->r0 + ((double)i - 0.5) * r->dr;
        }
        if ((st = upload_tab(r, &r->rC, rc))) return fail(st);
        if ((st = upload_tab(r, &r->rT, rt))) return fail(st);
    }
    if (r->cfg.direct_cols <= 0) r->cfg.direct_cols = vv::default_direct_cols(cfg->np);
    r->nbx = field_blocks(r);
    {
        const size_t np_ = 2 * (size_t)r->nbx + 3 * (size_t)r->nbx + 2 * (size_t)vv::MG_NORM_MAXB;
        if (hipMalloc(&r->part, np_ * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->part, 0, np_ * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
        r->fpart = r->part;
        r->upart = r->part + 2 * (size_t)r->nbx;
        r->npart = r->part + 5 * (size_t)r->nbx;
    }
    if (hipMalloc(&r->mg_state, 128) != hipSuccess) return fail(VV_ENOMEM);
    if (hipMemsetAsync(r->mg_state, 0, 128, r->stream) != hipSuccess) return fail(VV_EHIP);
    r->mg_flags = static_cast<int *>(r->mg_state);
    r->mg_bits = reinterpret_cast<unsigned long long *>(static_cast<char *>(r->mg_state) + 32);
    r->mg_relres = reinterpret_cast<double *>(static_cast<char *>(r->mg_state) + 64);
    try {
        if ((st = mg_build(r))) return fail(st);
    } catch (const std::bad_alloc &) {
        set_error("vv_create: host memory for the Poisson tables exhausted");
        return fail(VV_ENOMEM);
    }
    r->diag_cap = 1 << 16;
    if (hipMalloc(&r->diag, (size_t)r->diag_cap * 2 * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hipMalloc(&r->kepart, (vv::KE_MAXB + 1) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    if (nlev) *nlev = r->nlev_v;
    *out = r;
    return VV_OK;
}

int vv_destroy(vv_run *r)
{
    if (!r) return VV_OK;
    {   // the lock is released before the run is deleted
        std::lock_guard<std::recursive_mutex> lk(r->mu);
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        double *fields[] = {r->C, r->Cp, r->Cn, r->Cf, r->th, r->thp, r->thn, r->thf, r->uel, r->vel,
                            r->uelp, r->velp, r->vrh, r->urh, r->dCr, r->dCh, r->dthrC, r->dthhC,
                            r->dthrT, r->dthhT, r->mT2, r->mT, r->mU2, r->mV2, r->rC, r->rT, r->tab_r1, r->tab_rT,
                            r->tab_p, r->mg_pool, r->mg_ftw, r->mg_hist, r->diag, r->part, r->kepart};
        for (double *f : fields)
            if (f) (void)hipFree(f);
        if (r->mg_pool32) (void)hipFree(r->mg_pool32);
        if (r->mg_state) (void)hipFree(r->mg_state);
        for (hipEvent_t e : r->ev_pool) (void)hipEventDestroy(e);
        if (r->ev0) (void)hipEventDestroy(r->ev0);
        if (r->ev1) (void)hipEventDestroy(r->ev1);
        if (r->stream) (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return VV_OK;
}

int vv_reset(vv_run *r, const vv_config *cfg)
{
    if (!r || !cfg) { set_error("vv_reset: null argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    if (!vv::same_shape(r->cfg, *cfg)) {
        set_error("vv_reset: the configuration's grid, geometry or solver differs from the run's");
        return VV_EINVAL;
    }
    if (!(cfg->dt > 0) || !(cfg->Pr > 0) || (cfg->geom == VV_GEOM_ANNULUS && cfg->Re == 0.0)) {
        set_error("bad physical parameters");
        return VV_EINVAL;
    }
    vv_config c = *cfg;
    if (c.rtol <= 0) c.rtol = 1e-10;
    if (c.max_cycles <= 0) c.max_cycles = 100;
    if (c.robert < 0) c.robert = 0.02;
    c.nu1 = r->cfg.nu1;
    c.nu2 = r->cfg.nu2;
    c.direct_cols = r->cfg.direct_cols;
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipStreamSynchronize(r->stream));
    // every state buffer back to zero, as vv_create leaves it
    for (double *f : {r->C, r->Cp, r->Cn, r->Cf, r->th, r->thp, r->thn, r->thf, r->uel, r->vel, r->uelp, r->velp,
                      r->vrh, r->urh, r->dCr, r->dCh, r->dthrC, r->dthhC, r->dthrT, r->dthhT, r->mT2, r->mT, r->mU2,
                      r->mV2})
        VV_HIP(hipMemsetAsync(f, 0, r->fbytes, r->stream));
    for (int p = 0; p < r->mg.nprob; ++p) {
        const vv::MgProblemDev &P = r->mg.prob[p];
        for (int l = 0; l < P.nlev; ++l) {
            const vv::MgLevelDev &L = P.lev[l];
            const size_t cnt = (size_t)P.n * L.ncol;
            for (double *b : {L.u, L.f, L.u2, L.xm})
                if (b) VV_HIP(hipMemsetAsync(b, 0, cnt * sizeof(double), r->stream));
            if (L.dlt) VV_HIP(hipMemsetAsync(L.dlt, 0, cnt / 2 * sizeof(double), r->stream));
            for (float *b : {L.u32, L.u2_32, L.f32})
                if (b) VV_HIP(hipMemsetAsync(b, 0, cnt * sizeof(float), r->stream));
            if (L.dlt32) VV_HIP(hipMemsetAsync(L.dlt32, 0, cnt / 2 * sizeof(float), r->stream));
        }
    }
    VV_HIP(hipMemsetAsync(r->mg_state, 0, 128, r->stream));
    if (r->mg_hist) VV_HIP(hipMemsetAsync(r->mg_hist, 0, 2 * vv::MG_HIST * sizeof(double), r->stream));
    VV_HIP(hipMemsetAsync(r->diag, 0, (size_t)r->diag_cap * 2 * sizeof(double), r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    r->cfg = c;
    r->steps = 0;
    r->last_cycles[0] = r->last_cycles[1] = 0;
    r->last_relres[0] = r->last_relres[1] = 0.0;
    pending(r) = Pending{};
    return VV_OK;
}

int vv_put(vv_run *r, int field, const double *host, int ld)
{
    if (!r || !host || ld < r->nr + 2) { set_error("vv_put: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    double *d = field_ptr(r, field);
    if (!d) { set_error("vv_put: unknown field"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    const size_t w = (size_t)r->nr + 2;
    double *st = r->stage.get(w * r->ncols);
    if (!st) { set_error("vv_put: pinned staging buffer"); return VV_ENOMEM; }
    for (int c = 0; c < r->ncols; ++c) std::memcpy(st + c * w, host + (size_t)c * ld, w * sizeof(double));
    VV_HIP(hipMemcpy2DAsync(d, (size_t)r->ld * sizeof(double), st, w * sizeof(double), w * sizeof(double), r->ncols,
                            hipMemcpyHostToDevice, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vv_get(vv_run *r, int field, double *host, int ld)
{
    if (!r || !host || ld < r->nr + 2) { set_error("vv_get: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    double *d = field_ptr(r, field);
    if (!d) { set_error("vv_get: unknown field"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    const size_t w = (size_t)r->nr + 2;
    double *st = r->stage.get(w * r->ncols);
    if (!st) { set_error("vv_get: pinned staging buffer"); return VV_ENOMEM; }
    VV_HIP(hipMemcpy2DAsync(st, w * sizeof(double), d, (size_t)r->ld * sizeof(double), w * sizeof(double), r->ncols,
                            hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    for (int c = 0; c < r->ncols; ++c) std::memcpy(host + (size_t)c * ld, st + c * w, w * sizeof(double));
    return VV_OK;
}

int vv_start(vv_run *r)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipMemcpyAsync(r->Cp, r->C, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    VV_HIP(hipMemcpyAsync(r->thp, r->th, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    VV_HIP(hipMemcpyAsync(r->uelp, r->uel, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    VV_HIP(hipMemcpyAsync(r->velp, r->vel, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    r->steps = 0;
    r->last_cycles[0] = r->last_cycles[1] = 0;
    pending(r) = Pending{};
    return VV_OK;
}

int vv_cvorticity(vv_run *r, int first)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_compact_mask(r, 1 | 2 | 8 | 16);
    if (!st) st = launch_df(r, first ? 1 : 0, 1, 0);
    if (!st) {
        VV_HIP(hipStreamSynchronize(r->stream));
        pending(r).c = true;
    }
    return st;
}

int vv_temperature(vv_run *r, int first)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_compact_mask(r, 4 | 32);
    if (!st) st = launch_df(r, first ? 1 : 0, 0, 1);
    if (!st) {
        VV_HIP(hipStreamSynchronize(r->stream));
        pending(r).t = true;
    }
    return st;
}

int vv_heat(vv_run *r, double *trans1, double *trans2)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const double *th = pending(r).t ? r->thf : r->th;
    int st = launch_heat_from(r, th, r->diag);
    if (st) return st;
    double h[2];
    VV_HIP(hipMemcpyAsync(h, r->diag, 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = h[0];
    if (trans2) *trans2 = h[1];
    return VV_OK;
}

int vv_energy(vv_run *r, double *ke)
{
    if (!r || !ke) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int st = vv::launch_energy(r);
    if (st) return st;
    VV_HIP(hipMemcpyAsync(ke, r->kepart + vv::KE_MAXB, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vv_rhs(vv_run *r)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_rhs_guess(r, 3, 5);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv_poisson(vv_run *r, int which, int *cycles, double *relres)
{
    if (!r || (which != 2 && which != 3 && which != 5)) { set_error("vv_poisson: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int w = (which == 5) ? 3 : (which == 2 ? 1 : 2);
    int st = launch_rhs_fields(r, w, which);
    if (st) return st;
    int cyc[2] = {0, 0};
    double rel[2] = {0, 0};
    st = mg_solve(r, which, cyc, rel);
    if (st < 0) return st;
    int st2 = launch_mg_scatter_out(r, which);
    if (st2) return st2;
    VV_HIP(hipStreamSynchronize(r->stream));
    if (cycles) { cycles[0] = (which == 3) ? cyc[1] : cyc[0]; if (which == 5) cycles[1] = cyc[1]; }
    if (relres) { relres[0] = (which == 3) ? rel[1] : rel[0]; if (which == 5) relres[1] = rel[1]; }
    return st;
}

int vv_poisson_host(vv_run *r, int which, double *psi, double *psip, const double *rhs, int ld,
                    int *cycles, double *relres)
{
    if (!r || !psi || !psip || !rhs || (which != 2 && which != 3)) {
        set_error("vv_poisson_host: bad argument");
        return VV_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    int st;
    const int fpsi = (which == 2) ? VV_VEL : VV_UEL;
    const int fpsip = (which == 2) ? VV_VELP : VV_UELP;
    const int frhs = (which == 2) ? VV_VRH : VV_URH;
    if ((st = vv_put(r, fpsi, psi, ld))) return st;
    if ((st = vv_put(r, fpsip, psip, ld))) return st;
    if ((st = vv_put(r, frhs, rhs, ld))) return st;
    const int pst = vv_poisson(r, which, cycles, relres);
    if (pst < 0) return pst;
    if ((st = vv_get(r, fpsi, psi, ld))) return st;
    if ((st = vv_get(r, fpsip, psip, ld))) return st;
    return pst;
}

// Combines the per-workgroup partials of the last k_update: CFL maxima of u
// and v (NaN propagates) and the clamp count.
static int read_update_partials(vv_run *r, double cfl[2], int *nclamp)
{
    std::vector<double> h(3 * (size_t)r->nbx);
    VV_HIP(hipMemcpyAsync(h.data(), r->upart, h.size() * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    double cu = 0.0, cv = 0.0, nc = 0.0;
    for (int b = 0; b < r->nbx; ++b) {
        const double a = h[b], c = h[r->nbx + b];
        if (a > cu || a != a) cu = a;
        if (c > cv || c != c) cv = c;
        nc += h[2 * (size_t)r->nbx + b];
    }
    cfl[0] = cu;
    cfl[1] = cv;
    *nclamp = (int)nc;
    return VV_OK;
}

int vv_update(vv_run *r, int *nclamp)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    Pending &pd = pending(r);
    if (!pd.c) VV_HIP(hipMemcpyAsync(r->Cf, r->C, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    if (!pd.t) VV_HIP(hipMemcpyAsync(r->thf, r->th, r->fbytes, hipMemcpyDeviceToDevice, r->stream));
    int st = launch_update(r, (int)(r->steps + 1), 0);
    if (st) return st;
    double cfl[2];
    int nc = 0;
    if ((st = read_update_partials(r, cfl, &nc))) return st;
    rotate(r);
    pd = Pending{};
    r->steps += 1;
    if (nclamp) *nclamp = nc;
    return VV_OK;
}

int vv_solve_history(vv_run *r, double *relres_v, double *relres_u, int max, int *ncycles)
{
    if (!r || max < 0 || (max > 0 && (!relres_v || !relres_u))) { set_error("vv_solve_history: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipStreamSynchronize(r->stream));
    int cyc[2] = {0, 0};
    VV_HIP(hipMemcpy(cyc, r->mg.cycles, 2 * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> h(2 * vv::MG_HIST);
    VV_HIP(hipMemcpy(h.data(), r->mg_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < max; ++k) {
        relres_v[k] = (k < std::min(cyc[0], vv::MG_HIST)) ? h[k] : 0.0;
        relres_u[k] = (k < std::min(cyc[1], vv::MG_HIST)) ? h[vv::MG_HIST + k] : 0.0;
    }
    if (ncycles) {
        ncycles[0] = cyc[0];
        ncycles[1] = cyc[1];
    }
    return VV_OK;
}

int vv_step(vv_run *r, int nsteps, vv_step_info *info)
{
    if (!r || nsteps < 0) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    pending(r) = Pending{};
    int status = VV_OK;
    int cyc[2] = {0, 0};
    double rel[2] = {0, 0};
    // Deferred solves: with the whole plane direct, once a synchronised
    // solve has converged in one cycle, later steps queue that cycle and a
    // refinement cycle (empty when the first converged) with a device-side
    // check in the second's stop test (k_mg_stop) without waiting for the
    // host; the check's sticky word is read once at the end of the call.
    // Profiled runs and the V-cycle synchronise every solve.
    const bool direct_all = r->cfg.direct_cols >= r->np;
    bool any_deferred = false;
    VV_HIP(hipMemsetAsync(r->mg_flags + 4, 0, sizeof(int), r->stream));
    VV_HIP(hipEventRecord(r->ev0, r->stream));
    for (int s = 0; s < nsteps; ++s) {
        const int first = (r->steps == 0) ? 1 : 0;
        int st = launch_compact(r);
        if (!st) st = launch_df(r, first, 1, 1);
        if (!st) st = launch_heat(r, r->diag + 2 * (size_t)(r->steps % r->diag_cap));
        if (!st) st = launch_rhs_guess(r, 3, 5);
        if (st) return st;
        const bool defer = direct_all && !r->prof_on && r->last_cycles[0] == 1 && r->last_cycles[1] == 1;
        if (defer) {
            st = mg_enqueue_deferred(r);
            if (st) return st;
            any_deferred = true;
        } else {
            st = mg_solve(r, 5, cyc, rel);
            if (st < 0) return st;
            if (st == VV_NOT_CONVERGED) status = VV_NOT_CONVERGED;
        }
        if ((st = launch_update(r, (int)(r->steps + 1), 1))) return st;
        rotate(r);
        r->steps += 1;
        if (r->prof_on) {
            VV_HIP(hipStreamSynchronize(r->stream));
            if ((st = prof_resolve(r))) return st;
        }
    }
    VV_HIP(hipEventRecord(r->ev1, r->stream));
    VV_HIP(hipEventSynchronize(r->ev1));
    float ms = 0.f;
    VV_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    r->last_ms = ms;
    if (any_deferred) {   // the deferred solves' sticky check and the last solve's state
        struct { int flags[8]; unsigned long long bits[4]; double rel[2]; } h;
        VV_HIP(hipMemcpy(&h, r->mg_state, sizeof(h), hipMemcpyDeviceToHost));
        cyc[0] = h.flags[2];
        cyc[1] = h.flags[3];
        rel[0] = h.rel[0];
        rel[1] = h.rel[1];
        if (h.flags[4]) {
            r->last_cycles[0] = r->last_cycles[1] = 0;   // synchronise the next solve
            if (h.flags[4] & 2) {
                set_error("non-finite Poisson right-hand side");
                return VV_ENONFINITE;
            }
            status = VV_NOT_CONVERGED;
        }
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        if (r->steps > 0) {
            double tr[2];
            VV_HIP(hipMemcpy(tr, r->diag + 2 * (size_t)((r->steps - 1) % r->diag_cap),
                             2 * sizeof(double), hipMemcpyDeviceToHost));
            info->trans1 = tr[0];
            info->trans2 = tr[1];
        }
        double cfl[2];
        int nc = 0;
        int st = read_update_partials(r, cfl, &nc);
        if (st) return st;
        info->cflu = cfl[0];
        info->cflv = cfl[1];
        info->cycles_v = cyc[0];
        info->cycles_u = cyc[1];
        info->relres_v = rel[0];
        info->relres_u = rel[1];
        info->nclamp = nc;
        info->status = status;
    }
    return status;
}

int vv_heat_history(vv_run *r, double *trans, int n)
{
    if (!r || !trans || n < 0) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    long long avail = r->steps < r->diag_cap ? r->steps : r->diag_cap;
    long long cnt = n < avail ? n : avail;
    for (long long j = 0; j < cnt; ++j) {
        long long s = r->steps - cnt + j;
        VV_HIP(hipMemcpy(trans + 2 * j, r->diag + 2 * (size_t)(s % r->diag_cap), 2 * sizeof(double),
                         hipMemcpyDeviceToHost));
    }
    return (int)cnt >= 0 ? VV_OK : VV_EINVAL;
}

int vv_profile(vv_run *r, int enable)
{
    if (!r) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipStreamSynchronize(r->stream));
    int st = prof_resolve(r);
    if (st) return st;
    if (enable) {
        for (int k = 0; k < 16; ++k) { r->prof_launch[k] = 0; r->prof_ms[k] = 0; r->prof_bytes[k] = 0; }
    }
    r->prof_on = enable ? 1 : 0;
    return VV_OK;
}

int vv_profile_read(vv_run *r, int kernel, long long *launches, double *ms, double *bytes)
{
    if (!r || kernel < 0 || kernel >= VV_K_COUNT) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipStreamSynchronize(r->stream));
    int st = prof_resolve(r);
    if (st) return st;
    if (launches) *launches = r->prof_launch[kernel];
    if (ms) *ms = r->prof_ms[kernel];
    if (bytes) *bytes = r->prof_bytes[kernel];
    return VV_OK;
}

int vv_last_step_ms(vv_run *r, double *ms)
{
    if (!r || !ms) return VV_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    *ms = r->last_ms;
    return VV_OK;
}

}  // extern "C"
