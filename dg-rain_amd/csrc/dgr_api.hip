// dgr_api.hip -- the C ABI of libdgrain (include/dgrain.h): handle lifetime,
// element-centric neighbour tables, host<->device transposition between the
// reference layout (el*3+k) and the SoA planes, and the stage sequence.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/dgrain.h"
#include "dgr_device.h"

namespace dgr {
int launch_stage(int variant, int mode, bool wetdry, const MeshDev &m, const PhysDev &ph, const double *win,
                  double *wout, const double *wbase, const double *bzeta, const double *bqn, const int *wd,
                  double *means, double *fhat, Ctrl *ctrl, int slot, double *dbg_rhs, double *dbg_fhat,
                  hipStream_t s);
void launch_means(const MeshDev &m, const double *w, double *means, hipStream_t s);
void launch_geometry(const MeshDev &m, double g, double *geo, hipStream_t s);
void launch_node_minmax(const MeshDev &m, const double *means, double *nodemm, hipStream_t s);
void launch_limit(bool pdop, const MeshDev &m, const PhysDev &ph, double *w, const double *nodemm,
                  unsigned char *modify, Ctrl *ctrl, int stage, hipStream_t s);
void launch_step_mark(Ctrl *c, double dt, hipStream_t s);
void launch_pdop(const MeshDev &m, const PhysDev &ph, double *w, Ctrl *ctrl, hipStream_t s);
void launch_wd_status(const MeshDev &m, const PhysDev &ph, const double *w, int *wd, hipStream_t s);
void launch_step_begin(const PhysDev &ph, Ctrl *c, hipStream_t s);
void launch_step_end(const PhysDev &ph, Ctrl *c, int nparts, hipStream_t s);
void launch_lam_reduce(const PhysDev &ph, Ctrl *c, int slot, int nparts, hipStream_t s);
void launch_reset_maxlam(Ctrl *c, hipStream_t s);
void launch_math_selftest(int n, const double *x, const double *y, double *out, hipStream_t s);
}  // namespace dgr

using namespace dgr;

struct dgr_handle {
    int ne, nn, ned, pitch;
    int device;
    hipStream_t stream;
    dgr_params prm;
    PhysDev ph;
    MeshDev m;
    // device buffers
    double4 *d_nodes;
    int *d_ev, *d_nb, *d_info, *d_eid, *d_n2e_off, *d_n2e;
    double *d_bzeta, *d_bqn;
    double *d_w, *d_w1, *d_means, *d_nodemm;
    double *d_dbg;   // 18 planes: rhs (9) then fhat (9)
    double *d_geo;   // GX_COUNT planes of per-element geometry
    double *d_fhat;  // 9 planes of edge fluxes (split kernel variant)
    int variant;     // bit 0: geometry planes, bit 1: split flux/RHS kernels, 4/5: edge fluxes
    int4 *d_edges;   // {el_in, el_ex, va, vb} per edge
    int *d_einfo;    // packed edge facts (ED_*/EI_* bits)
    unsigned long long *d_lamparts;   // [2][maxparts] per-block max |lambda|
    int nparts[2];   // partials written by the last flux launch of each stage
    int *d_wd;
    unsigned char *d_modify;   // [2][pitch]
    Ctrl *d_ctrl;
    // host staging
    std::vector<double> h_planes;
    // optional HIP-event timing of the flux/RHS/update kernel (k_stage)
    int profiling;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used;
};

#define DGR_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "libdgrain: %s failed: %s\n", #call, hipGetErrorString(e_)); \
            return DGR_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

extern "C" void dgr_params_default(dgr_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->g = 9.81;
    p->h0 = 1e-3;
    p->velzero = 1e-2;
    p->cfl = 0.1;
    p->dt0 = 1e-7;
    p->final_time = HUGE_VAL;
    p->rain = 0.0;
    p->wetdry = 0;
    p->tide_code = 0;
    p->tide_period = 43200.0;
    p->record_modify = 1;
}

extern "C" const char *dgr_status_string(dgr_status s)
{
    switch (s) {
    case DGR_OK: return "ok";
    case DGR_ERR_ARG: return "invalid argument";
    case DGR_ERR_MESH: return "inconsistent mesh";
    case DGR_ERR_HIP: return "HIP runtime error";
    case DGR_ERR_NAN: return "NaN flux or wet/dry failure";
    case DGR_ERR_NOMEM: return "out of memory";
    case DGR_ERR_NODEV: return "no HIP device";
    case DGR_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown";
}

// local edge of element vertices ev[3] matching (v0,v1) (compute2DL.c:45-67)
static int local_edge(const int *ev, int v0, int v1)
{
    if ((v0 == ev[0] && v1 == ev[1]) || (v0 == ev[1] && v1 == ev[0])) return 0;
    if ((v0 == ev[1] && v1 == ev[2]) || (v0 == ev[2] && v1 == ev[1])) return 1;
    if ((v0 == ev[2] && v1 == ev[0]) || (v0 == ev[0] && v1 == ev[2])) return 2;
    return -1;
}

static dgr_status free_handle(dgr_handle *h)
{
    if (!h) return DGR_OK;
    void *ptrs[] = {h->d_nodes, h->d_ev, h->d_nb, h->d_info, h->d_eid, h->d_n2e_off, h->d_n2e, h->d_bzeta,
                    h->d_bqn, h->d_w, h->d_w1, h->d_means, h->d_nodemm, h->d_dbg, h->d_geo, h->d_fhat, h->d_wd, h->d_modify,
                    h->d_ctrl, h->d_edges, h->d_einfo, h->d_lamparts};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DGR_OK;
}

template <typename T>
static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

extern "C" dgr_status dgr_create(const dgr_mesh_view *mv, const dgr_params *prm, dgr_handle **out)
{
    if (!mv || !prm || !out) return DGR_ERR_ARG;
    *out = nullptr;
    const int ne = mv->num_el, nn = mv->num_nodes, ned = mv->num_edges;
    if (ne <= 0 || nn <= 0 || ned <= 0 || !mv->x || !mv->y || !mv->z || !mv->n || !mv->elto_vert ||
        !mv->edg_to_vert || !mv->edg_to_els || !mv->bdry)
        return DGR_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DGR_ERR_NODEV;

    // element-centric tables, one slot per (element, local edge)
    const int pitch = (ne + 63) / 64 * 64;
    std::vector<int> ev(3 * (size_t)pitch, 0), nb(3 * (size_t)pitch, 0), info(3 * (size_t)pitch, 0),
        eid(3 * (size_t)pitch, 0);
    std::vector<unsigned char> seen(3 * (size_t)ne, 0);
    std::vector<int4> edges(ned);
    std::vector<int> einfo(ned);
    for (int el = 0; el < ne; ++el)
        for (int k = 0; k < 3; ++k) {
            const int v = mv->elto_vert[3 * el + k];
            if (v < 0 || v >= nn) return DGR_ERR_MESH;
            ev[(size_t)k * pitch + el] = v;
        }
    const double *X = mv->x, *Y = mv->y;
    for (int i = 0; i < ned; ++i) {
        const int el_in = mv->edg_to_els[2 * i], el_ex = mv->edg_to_els[2 * i + 1];
        const int v0 = mv->edg_to_vert[2 * i], v1 = mv->edg_to_vert[2 * i + 1];
        if (el_in < 0 || el_in >= ne || el_ex < 0 || el_ex >= ne) return DGR_ERR_MESH;
        const int *ein = &mv->elto_vert[3 * el_in], *eex = &mv->elto_vert[3 * el_ex];
        const int kin = local_edge(ein, v0, v1), kex = local_edge(eex, v0, v1);
        if (kin < 0 || kex < 0) return DGR_ERR_MESH;
        const double jac = X[ein[0]] * (Y[ein[1]] - Y[ein[2]]) + X[ein[1]] * (Y[ein[2]] - Y[ein[0]]) +
                           X[ein[2]] * (Y[ein[0]] - Y[ein[1]]);   // compute2DL.c:100-102
        const int sg = jac < 0 ? -1 : (jac == 0 ? 0 : 1);
        const int bp = mv->bdry[i];
        const bool boundary = el_in == el_ex;
        int common = ((sg + 1) << EI_SGN_SHIFT) | ((bp < 0 ? 0 : (bp > 255 ? 255 : bp)) << EI_CODE_SHIFT);
        if (bp) common |= EI_PRESCRIBED;
        if (boundary) common |= EI_BOUNDARY;
        edges[i] = make_int4(el_in, el_ex, ein[kin], ein[(kin + 1) % 3]);
        einfo[i] = common | kin | (kex << ED_KEX_SHIFT);
        // el_in slot
        {
            size_t s = (size_t)kin * pitch + el_in;
            if (seen[3 * (size_t)el_in + kin]++) return DGR_ERR_MESH;
            nb[s] = el_ex;
            info[s] = common | EI_IS_IN | EI_DIR_SAME | (kex & EI_KNB_MASK);
            eid[s] = i;
        }
        if (!boundary) {
            size_t s = (size_t)kex * pitch + el_ex;
            if (seen[3 * (size_t)el_ex + kex]++) return DGR_ERR_MESH;
            const bool same = ein[kin] == eex[kex];
            nb[s] = el_in;
            info[s] = common | (same ? EI_DIR_SAME : 0) | (kin & EI_KNB_MASK);
            eid[s] = i;
        }
    }
    for (size_t i = 0; i < seen.size(); ++i)
        if (seen[i] != 1) return DGR_ERR_MESH;   // every Fhat slot written exactly once
    // node -> element CSR, ascending element order (SlopeLimiter2D.c:73-89)
    std::vector<int> off(nn + 1, 0), n2e(3 * (size_t)ne);
    for (int i = 0; i < 3 * ne; ++i) off[mv->elto_vert[i] + 1]++;
    for (int i = 0; i < nn; ++i) off[i + 1] += off[i];
    {
        std::vector<int> fill(nn, 0);
        for (int el = 0; el < ne; ++el)
            for (int k = 0; k < 3; ++k) {
                const int v = mv->elto_vert[3 * el + k];
                n2e[off[v] + fill[v]++] = el;
            }
    }
    std::vector<double4> nodes(nn);
    for (int i = 0; i < nn; ++i) nodes[i] = make_double4(X[i], Y[i], mv->z[i], mv->n[i]);
    // minEdgLength (extract_mesh_attributes.c:297-303)
    double mel = 9999999;
    for (int e = 0; e < ned; ++e) {
        const int n0 = mv->edg_to_vert[2 * e], n1 = mv->edg_to_vert[2 * e + 1];
        const double len = std::sqrt((X[n1] - X[n0]) * (X[n1] - X[n0]) + (Y[n1] - Y[n0]) * (Y[n1] - Y[n0]));
        mel = (mel < len) ? mel : len;
    }

    dgr_handle *h = new dgr_handle();   // value-initialised: all pointers null
    h->ne = ne; h->nn = nn; h->ned = ned; h->pitch = pitch;
    h->prm = *prm;
    (void)hipGetDevice(&h->device);
#define DGR_TRY(call)                                          \
    do {                                                       \
        if ((call) != hipSuccess) { free_handle(h); return DGR_ERR_HIP; } \
    } while (0)
    DGR_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const size_t P = pitch;
    DGR_TRY(dalloc(&h->d_nodes, nn));
    DGR_TRY(dalloc(&h->d_ev, 3 * P));
    DGR_TRY(dalloc(&h->d_nb, 3 * P));
    DGR_TRY(dalloc(&h->d_info, 3 * P));
    DGR_TRY(dalloc(&h->d_eid, 3 * P));
    DGR_TRY(dalloc(&h->d_n2e_off, nn + 1));
    DGR_TRY(dalloc(&h->d_n2e, 3 * (size_t)ne));
    DGR_TRY(dalloc(&h->d_bzeta, ned));
    DGR_TRY(dalloc(&h->d_bqn, ned));
    DGR_TRY(dalloc(&h->d_w, 9 * P));
    DGR_TRY(dalloc(&h->d_w1, 9 * P));
    DGR_TRY(dalloc(&h->d_means, 3 * P));
    DGR_TRY(dalloc(&h->d_nodemm, 6 * (size_t)nn));
    DGR_TRY(dalloc(&h->d_wd, P));
    DGR_TRY(dalloc(&h->d_modify, 2 * P));
    DGR_TRY(dalloc(&h->d_ctrl, 1));
    DGR_TRY(dalloc(&h->d_geo, (size_t)GX_COUNT * P));
    DGR_TRY(dalloc(&h->d_fhat, 9 * P));
    DGR_TRY(dalloc(&h->d_edges, ned));
    DGR_TRY(dalloc(&h->d_einfo, ned));
    const int maxparts = ((ne > ned ? ne : ned) + 255) / 256;
    DGR_TRY(dalloc(&h->d_lamparts, 2 * (size_t)maxparts));
    hipStream_t s = h->stream;
    DGR_TRY(hipMemcpyAsync(h->d_nodes, nodes.data(), nn * sizeof(double4), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_ev, ev.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_nb, nb.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_info, info.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_eid, eid.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_edges, edges.data(), ned * sizeof(int4), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_einfo, einfo.data(), ned * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_n2e_off, off.data(), (nn + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_n2e, n2e.data(), 3 * (size_t)ne * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemsetAsync(h->d_w, 0, 9 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_w1, 0, 9 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_modify, 0, 2 * P, s));
    DGR_TRY(hipMemsetAsync(h->d_lamparts, 0, 2 * (size_t)maxparts * sizeof(unsigned long long), s));
    {
        std::vector<int> wd(P, -1);   // initialize.c:126
        DGR_TRY(hipMemcpyAsync(h->d_wd, wd.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
    }
    {
        // boundary_conditions.c:10-83 defaults
        std::vector<double> bz(ned, 0.0), bq(ned, 0.0);
        for (int e = 0; e < ned; ++e) {
            const int bp = mv->bdry[e];
            if (bp == 1) { bq[e] = -0.042 / 0.914; bz[e] = -1000; }
            if (bp == 2) { bz[e] = 0.296; bq[e] = 1000000; }
            if (bp == 3) { bz[e] = -10000; bq[e] = -0.1275 / 0.914; }
        }
        DGR_TRY(hipMemcpyAsync(h->d_bzeta, bz.data(), ned * sizeof(double), hipMemcpyHostToDevice, s));
        DGR_TRY(hipMemcpyAsync(h->d_bqn, bq.data(), ned * sizeof(double), hipMemcpyHostToDevice, s));
    }
    {
        Ctrl c;
        memset(&c, 0, sizeof(c));
        c.t = 0.0;
        c.dt = prm->dt0;
        c.err_key = ~0ULL;
        c.nan_pos[0] = c.nan_pos[1] = -1;
        DGR_TRY(hipMemcpyAsync(h->d_ctrl, &c, sizeof(c), hipMemcpyHostToDevice, s));
    }
    DGR_TRY(hipStreamSynchronize(s));
#undef DGR_TRY
    h->m.ne = ne; h->m.nn = nn; h->m.pitch = pitch;
    h->m.nodes = h->d_nodes; h->m.ev = h->d_ev; h->m.nb = h->d_nb; h->m.info = h->d_info; h->m.eid = h->d_eid;
    h->m.n2e_off = h->d_n2e_off; h->m.n2e = h->d_n2e;
    h->m.geo = h->d_geo;
    h->m.ned = ned; h->m.edges = h->d_edges; h->m.einfo = h->d_einfo;
    launch_geometry(h->m, prm->g, h->d_geo, s);
    PhysDev &ph = h->ph;
    ph.g = prm->g; ph.h0 = prm->h0; ph.velzero = prm->velzero; ph.rain = prm->rain;
    ph.cfl = prm->cfl; ph.min_edge = mel; ph.final_time = prm->final_time;
    ph.tide_datum = prm->tide_datum; ph.tide_amp = prm->tide_amp; ph.tide_period = prm->tide_period;
    ph.tide_phase = prm->tide_phase;
    ph.wetdry = prm->wetdry; ph.tide_code = prm->tide_code;
    ph.lamparts = h->d_lamparts; ph.maxparts = maxparts;
    h->h_planes.resize(9 * P);
    // tide value for t = 0
    launch_step_begin(h->ph, h->d_ctrl, s);
    if (hipStreamSynchronize(s) != hipSuccess) { free_handle(h); return DGR_ERR_HIP; }
    *out = h;
    return DGR_OK;
}

extern "C" dgr_status dgr_destroy(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    (void)hipStreamSynchronize(h->stream);
    return free_handle(h);
}

static dgr_status planes_up(dgr_handle *h, double *dst, const double *a, const double *b, const double *c)
{
    const size_t P = h->pitch;
    const double *src[3] = {a, b, c};
    for (int f = 0; f < 3; ++f)
        for (int el = 0; el < h->ne; ++el)
            for (int k = 0; k < 3; ++k) h->h_planes[(size_t)(3 * f + k) * P + el] = src[f][3 * el + k];
    DGR_HIP(hipMemcpyAsync(dst, h->h_planes.data(), 9 * P * sizeof(double), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

static dgr_status planes_down(dgr_handle *h, const double *srcd, double *a, double *b, double *c)
{
    const size_t P = h->pitch;
    DGR_HIP(hipMemcpyAsync(h->h_planes.data(), srcd, 9 * P * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    double *dst[3] = {a, b, c};
    for (int f = 0; f < 3; ++f)
        if (dst[f])
            for (int el = 0; el < h->ne; ++el)
                for (int k = 0; k < 3; ++k) dst[f][3 * el + k] = h->h_planes[(size_t)(3 * f + k) * P + el];
    return DGR_OK;
}

extern "C" dgr_status dgr_set_state(dgr_handle *h, const double *z, const double *qx, const double *qy)
{
    if (!h || !z || !qx || !qy) return DGR_ERR_ARG;
    const unsigned long long none = ~0ULL;   // a new state clears a recorded error
    DGR_HIP(hipMemcpyAsync(&h->d_ctrl->err_key, &none, sizeof(none), hipMemcpyHostToDevice, h->stream));
    return planes_up(h, h->d_w, z, qx, qy);
}

extern "C" dgr_status dgr_get_state(dgr_handle *h, double *z, double *qx, double *qy)
{
    if (!h) return DGR_ERR_ARG;
    return planes_down(h, h->d_w, z, qx, qy);
}

extern "C" dgr_status dgr_set_wd(dgr_handle *h, const int *wd)
{
    if (!h || !wd) return DGR_ERR_ARG;
    DGR_HIP(hipMemcpyAsync(h->d_wd, wd, h->ne * sizeof(int), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_get_wd(dgr_handle *h, int *wd)
{
    if (!h || !wd) return DGR_ERR_ARG;
    DGR_HIP(hipMemcpyAsync(wd, h->d_wd, h->ne * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_get_modify(dgr_handle *h, int stage, unsigned char *mod)
{
    if (!h || !mod || stage < 0 || stage > 1) return DGR_ERR_ARG;
    DGR_HIP(hipMemcpyAsync(mod, h->d_modify + (size_t)stage * h->pitch, h->ne, hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_set_boundary(dgr_handle *h, const double *bz, const double *bq)
{
    if (!h || !bz || !bq) return DGR_ERR_ARG;
    DGR_HIP(hipMemcpyAsync(h->d_bzeta, bz, h->ned * sizeof(double), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(h->d_bqn, bq, h->ned * sizeof(double), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_set_time(dgr_handle *h, double t, double dt)
{
    if (!h) return DGR_ERR_ARG;
    double v[2] = {t, dt};
    DGR_HIP(hipMemcpyAsync(h->d_ctrl, v, sizeof(v), hipMemcpyHostToDevice, h->stream));
    launch_step_begin(h->ph, h->d_ctrl, h->stream);   // tide value at t
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

static dgr_status read_ctrl(dgr_handle *h, Ctrl *c)
{
    DGR_HIP(hipMemcpyAsync(c, h->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

static inline unsigned char *modify_ptr(dgr_handle *h, int stage)
{
    return h->prm.record_modify ? h->d_modify + (size_t)stage * h->pitch : nullptr;
}

extern "C" dgr_status dgr_compute_rhs(dgr_handle *h, double *rz, double *rqx, double *rqy, double *f1, double *f2,
                                      double *f3, double *max_lambda)
{
    if (!h) return DGR_ERR_ARG;
    if (!h->d_dbg) DGR_HIP(dalloc(&h->d_dbg, 18 * (size_t)h->pitch));
    launch_reset_maxlam(h->d_ctrl, h->stream);
    h->nparts[0] = launch_stage(h->variant, 0, h->prm.wetdry != 0, h->m, h->ph, h->d_w, nullptr, nullptr, h->d_bzeta,
                                h->d_bqn, h->d_wd, nullptr, h->d_fhat, h->d_ctrl, 0, h->d_dbg,
                                h->d_dbg + 9 * (size_t)h->pitch, h->stream);
    launch_lam_reduce(h->ph, h->d_ctrl, 0, h->nparts[0], h->stream);
    launch_stage(h->variant, 3, h->prm.wetdry != 0, h->m, h->ph, h->d_w, nullptr, nullptr, h->d_bzeta, h->d_bqn,
                 h->d_wd, nullptr, h->d_fhat, h->d_ctrl, 0, nullptr, nullptr, h->stream);
    DGR_HIP(hipGetLastError());
    if (rz || rqx || rqy) {
        dgr_status st = planes_down(h, h->d_dbg, rz, rqx, rqy);
        if (st) return st;
    }
    if (f1 || f2 || f3) {
        dgr_status st = planes_down(h, h->d_dbg + 9 * (size_t)h->pitch, f1, f2, f3);
        if (st) return st;
    }
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    if (max_lambda) *max_lambda = maxlam_of(c, 0);
    return err_code_of(c.err_key) ? DGR_ERR_NAN : DGR_OK;
}

extern "C" dgr_status dgr_limit(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_means(h->m, h->d_w, h->d_means, h->stream);
    launch_node_minmax(h->m, h->d_means, h->d_nodemm, h->stream);
    launch_limit(false, h->m, h->ph, h->d_w, h->d_nodemm, modify_ptr(h, 0), h->d_ctrl, 0, h->stream);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_wetdry_status(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_wd_status(h->m, h->ph, h->d_w, h->d_wd, h->stream);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_pdop(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_pdop(h->m, h->ph, h->d_w, h->d_ctrl, h->stream);
    DGR_HIP(hipGetLastError());
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    return err_code_of(c.err_key) ? DGR_ERR_NAN : DGR_OK;
}

// event pairs around k_stage launches when profiling is on
static void mark(dgr_handle *h)
{
    if (!h->profiling) return;
    if (h->ev_used == h->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        h->ev_pool.push_back(e);
    }
    (void)hipEventRecord(h->ev_pool[h->ev_used++], h->stream);
}

// the two stages of oneTimeStep2D (time_evolution.c:10-246; wet/dry calls at the
// places ChanNet/time_evolution.c:218-427 puts them), enqueued on h->stream
static void enqueue_stages(dgr_handle *h)
{
    const bool wdon = h->prm.wetdry != 0;
    hipStream_t s = h->stream;
    if (wdon) launch_wd_status(h->m, h->ph, h->d_w, h->d_wd, s);
    mark(h);
    h->nparts[0] = launch_stage(h->variant, 1, wdon, h->m, h->ph, h->d_w, h->d_w1, nullptr, h->d_bzeta, h->d_bqn,
                                h->d_wd, h->d_means, h->d_fhat, h->d_ctrl, 0, nullptr, nullptr, s);
    mark(h);
    launch_node_minmax(h->m, h->d_means, h->d_nodemm, s);
    launch_limit(wdon, h->m, h->ph, h->d_w1, h->d_nodemm, modify_ptr(h, 0), h->d_ctrl, 0, s);
    if (wdon) launch_wd_status(h->m, h->ph, h->d_w1, h->d_wd, s);
    mark(h);
    h->nparts[1] = launch_stage(h->variant, 2, wdon, h->m, h->ph, h->d_w1, h->d_w, h->d_w, h->d_bzeta, h->d_bqn,
                                h->d_wd, h->d_means, h->d_fhat, h->d_ctrl, 1, nullptr, nullptr, s);
    mark(h);
    // stage 2's max |lambda| sets the next dt; repair its fold if a NaN operand was seen
    launch_stage(h->variant, 3, wdon, h->m, h->ph, h->d_w1, nullptr, nullptr, h->d_bzeta, h->d_bqn, h->d_wd,
                 nullptr, h->d_fhat, h->d_ctrl, 1, nullptr, nullptr, s);
    launch_node_minmax(h->m, h->d_means, h->d_nodemm, s);
    launch_limit(wdon, h->m, h->ph, h->d_w, h->d_nodemm, modify_ptr(h, 1), h->d_ctrl, 1, s);
}

extern "C" dgr_status dgr_one_time_step(dgr_handle *h, double t, double dt)
{
    if (!h) return DGR_ERR_ARG;
    (void)t;
    launch_step_mark(h->d_ctrl, dt, h->stream);
    enqueue_stages(h);
    launch_lam_reduce(h->ph, h->d_ctrl, 1, h->nparts[1], h->stream);
    DGR_HIP(hipGetLastError());
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    return err_code_of(c.err_key) ? DGR_ERR_NAN : DGR_OK;
}

extern "C" dgr_status dgr_run(dgr_handle *h, int nsteps, dgr_report *rep)
{
    if (!h || nsteps < 0) return DGR_ERR_ARG;
    for (int n = 0; n < nsteps; ++n) {
        launch_step_begin(h->ph, h->d_ctrl, h->stream);
        enqueue_stages(h);
        launch_step_end(h->ph, h->d_ctrl, h->nparts[1], h->stream);
    }
    DGR_HIP(hipGetLastError());
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    if (rep) {
        rep->t = c.t;
        rep->dt = c.dt;
        rep->max_lambda = maxlam_of(c, 1);
        rep->error_code = err_code_of(c.err_key);
        rep->error_where = err_where_of(c.err_key);
    }
    return err_code_of(c.err_key) ? DGR_ERR_NAN : DGR_OK;
}

extern "C" dgr_status dgr_get_report(dgr_handle *h, dgr_report *rep)
{
    if (!h || !rep) return DGR_ERR_ARG;
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    rep->t = c.t;
    rep->dt = c.dt;
    rep->max_lambda = maxlam_of(c, 1);
    rep->error_code = err_code_of(c.err_key);
    rep->error_where = err_where_of(c.err_key);
    return rep->error_code ? DGR_ERR_NAN : DGR_OK;
}

extern "C" dgr_status dgr_set_final_time(dgr_handle *h, double final_time)
{
    if (!h) return DGR_ERR_ARG;
    h->prm.final_time = final_time;
    h->ph.final_time = final_time;
    return DGR_OK;
}

extern "C" dgr_status dgr_last_error(dgr_handle *h, int *code, int *where)
{
    if (!h) return DGR_ERR_ARG;
    Ctrl c;
    dgr_status st = read_ctrl(h, &c);
    if (st) return st;
    if (code) *code = err_code_of(c.err_key);
    if (where) *where = err_where_of(c.err_key);
    return DGR_OK;
}

extern "C" dgr_status dgr_sync(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_device_state(dgr_handle *h, double **planes, int *pitch)
{
    if (!h || !planes || !pitch) return DGR_ERR_ARG;
    *planes = h->d_w;
    *pitch = h->pitch;
    return DGR_OK;
}

extern "C" dgr_status dgr_math_selftest(int n, const double *x, const double *y, double *out)
{
    if (n <= 0 || !x || !y || !out) return DGR_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DGR_ERR_NODEV;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    DGR_HIP(dalloc(&dx, n));
    DGR_HIP(dalloc(&dy, n));
    DGR_HIP(dalloc(&dout, 5 * (size_t)n));
    DGR_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
    DGR_HIP(hipMemcpy(dy, y, n * sizeof(double), hipMemcpyHostToDevice));
    launch_math_selftest(n, dx, dy, dout, 0);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipMemcpy(out, dout, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return DGR_OK;
}

extern "C" dgr_status dgr_profile_begin(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    h->profiling = 1;
    h->ev_used = 0;
    return DGR_OK;
}

extern "C" dgr_status dgr_profile_end(dgr_handle *h, double *total_ms, int *launches)
{
    if (!h || !total_ms || !launches) return DGR_ERR_ARG;
    DGR_HIP(hipStreamSynchronize(h->stream));
    double tot = 0.0;
    int n = 0;
    for (size_t i = 0; i + 1 < h->ev_used; i += 2) {
        float ms = 0.f;
        DGR_HIP(hipEventElapsedTime(&ms, h->ev_pool[i], h->ev_pool[i + 1]));
        tot += ms;
        ++n;
    }
    h->profiling = 0;
    h->ev_used = 0;
    *total_ms = tot;
    *launches = n;
    return DGR_OK;
}

extern "C" dgr_status dgr_set_variant(dgr_handle *h, int variant)
{
    if (!h || variant < 0 || variant >= DGR_NUM_VARIANTS) return DGR_ERR_ARG;
    h->variant = variant;
    return DGR_OK;
}

extern "C" dgr_status dgr_get_variant(dgr_handle *h, int *variant)
{
    if (!h || !variant) return DGR_ERR_ARG;
    *variant = h->variant;
    return DGR_OK;
}

// Times `steps` steps of every kernel variant from the current state and keeps
// the fastest.  All variants compute the same bits; the state, WD flags and
// step bookkeeping are restored afterwards.
extern "C" dgr_status dgr_autotune(dgr_handle *h, int steps, int *chosen, double *ms_per_step)
{
    if (!h || steps <= 0) return DGR_ERR_ARG;
    const size_t P = h->pitch;
    double *w_save = nullptr;
    int *wd_save = nullptr;
    Ctrl c_save;
    DGR_HIP(dalloc(&w_save, 9 * P));
    DGR_HIP(dalloc(&wd_save, P));
    DGR_HIP(hipMemcpyAsync(w_save, h->d_w, 9 * P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(wd_save, h->d_wd, P * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(&c_save, h->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    hipEvent_t e0, e1;
    DGR_HIP(hipEventCreate(&e0));
    DGR_HIP(hipEventCreate(&e1));
    int best = h->variant;
    float best_ms = 1e30f;
    for (int v = 0; v < DGR_NUM_VARIANTS; ++v) {
        h->variant = v;
        // one untimed step to warm the code object, then the timed steps
        launch_step_begin(h->ph, h->d_ctrl, h->stream);
        enqueue_stages(h);
        launch_step_end(h->ph, h->d_ctrl, h->nparts[1], h->stream);
        DGR_HIP(hipEventRecord(e0, h->stream));
        for (int n = 0; n < steps; ++n) {
            launch_step_begin(h->ph, h->d_ctrl, h->stream);
            enqueue_stages(h);
            launch_step_end(h->ph, h->d_ctrl, h->nparts[1], h->stream);
        }
        DGR_HIP(hipEventRecord(e1, h->stream));
        DGR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DGR_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms_per_step) ms_per_step[v] = ms / steps;
        if (ms < best_ms) { best_ms = ms; best = v; }
        DGR_HIP(hipMemcpyAsync(h->d_w, w_save, 9 * P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        DGR_HIP(hipMemcpyAsync(h->d_wd, wd_save, P * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
        DGR_HIP(hipMemcpyAsync(h->d_ctrl, &c_save, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
        DGR_HIP(hipStreamSynchronize(h->stream));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(w_save);
    (void)hipFree(wd_save);
    h->variant = best;
    if (chosen) *chosen = best;
    return DGR_OK;
}
