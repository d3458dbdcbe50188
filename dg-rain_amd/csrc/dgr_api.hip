// dgr_api.hip -- the C ABI of libdgrain (include/dgrain.h): handle lifetime,
// element-centric and edge tables of one or more members, host<->device
// transposition between the reference layout (el*3+k) and the SoA planes, the
// stage sequence and its hipGraph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/dgrain.h"
#include "dgr_device.h"

#include "dgr_handle.h"

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
    p->bc_sentinels = 1;
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
    void *ptrs[] = {h->d_nodes, h->d_zn, h->d_ev, h->d_nb, h->d_info, h->d_eid, h->d_n2e_off, h->d_n2e, h->d_nl_off, h->d_nl, h->d_loc_off, h->d_locw, h->d_edges,
                    h->d_einfo, h->d_egeo, h->d_mem, h->d_eblk_mem, h->d_dblk_mem, h->d_bzeta, h->d_bqn, h->d_w, h->d_w1,
                    h->d_means, h->d_nodemm, h->d_dbg, h->d_fhat, h->d_wd, h->d_modify, h->d_ctrl, h->d_lamparts,
                    h->d_nanparts, h->d_wdn, h->d_hydro, h->d_tape, h->d_tiemax, h->d_tiecnt,
                    h->d_cpl_spec, h->d_cpl_vals};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DGR_OK;
}

static inline int pad256(int n) { return (n + 255) / 256 * 256; }

// Host tables of all members, in the concatenated (padded) index space.
struct HostTables {
    std::vector<int> ev, nb, info, eid, off, n2e, einfo, eblk_mem, dblk_mem;
    std::vector<int> nl_off, nl, loc_off;
    std::vector<unsigned> locw;
    std::vector<int4> edges;
    std::vector<double4> nodes;
    std::vector<double> bz, bq;
    std::vector<MemberDev> mdev;
};

// Validates member `mi` and writes its tables at its offsets.
static dgr_status build_member(const dgr_mesh_view *mv, const MemberHost &M, int mi, int pitch, HostTables &T)
{
    const int ne = M.ne, nn = M.nn, ned = M.ned;
    const int eo = M.el_off, no = M.nd_off, dof = M.ed_off;
    for (int el = 0; el < ne; ++el)
        for (int k = 0; k < 3; ++k) {
            const int v = mv->elto_vert[3 * el + k];
            if (v < 0 || v >= nn) return DGR_ERR_MESH;
            T.ev[(size_t)k * pitch + eo + el] = v + no;
        }
    std::vector<unsigned char> seen(3 * (size_t)ne, 0);
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
        T.edges[dof + i] = make_int4(eo + el_in, eo + el_ex, no + ein[kin], no + ein[(kin + 1) % 3]);
        T.einfo[dof + i] = common | kin | (kex << ED_KEX_SHIFT);
        {
            const size_t s = (size_t)kin * pitch + eo + el_in;
            if (seen[3 * (size_t)el_in + kin]++) return DGR_ERR_MESH;
            T.nb[s] = eo + el_ex;
            T.info[s] = common | EI_IS_IN | EI_DIR_SAME | (kex & EI_KNB_MASK);
            T.eid[s] = dof + i;
        }
        if (!boundary) {
            const size_t s = (size_t)kex * pitch + eo + el_ex;
            if (seen[3 * (size_t)el_ex + kex]++) return DGR_ERR_MESH;
            const bool same = ein[kin] == eex[kex];
            T.nb[s] = eo + el_in;
            T.info[s] = common | (same ? EI_DIR_SAME : 0) | (kin & EI_KNB_MASK);
            T.eid[s] = dof + i;
        }
        // boundary_conditions.c:10-83 defaults
        if (bp == 1) { T.bq[dof + i] = -0.042 / 0.914; T.bz[dof + i] = -1000; }
        if (bp == 2) { T.bz[dof + i] = 0.296; T.bq[dof + i] = 1000000; }
        if (bp == 3) { T.bz[dof + i] = -10000; T.bq[dof + i] = -0.1275 / 0.914; }
    }
    for (size_t i = 0; i < seen.size(); ++i)
        if (seen[i] != 1) return DGR_ERR_MESH;   // every Fhat slot written exactly once
    // node -> element CSR, ascending element order (SlopeLimiter2D.c:73-89);
    // each entry also carries the node's local vertex index in the element
    // (bits 30-31, see N2E_EL_MASK)
    {
        std::vector<int> cnt(nn + 1, 0);
        for (int i = 0; i < 3 * ne; ++i) cnt[mv->elto_vert[i] + 1]++;
        for (int i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
        const int base = T.off[no];   // CSR entries of earlier members
        for (int i = 0; i <= nn; ++i) T.off[no + i] = base + cnt[i];
        std::vector<int> fill(nn, 0);
        for (int el = 0; el < ne; ++el)
            for (int k = 0; k < 3; ++k) {
                const int v = mv->elto_vert[3 * el + k];
                T.n2e[base + cnt[v] + fill[v]++] = (eo + el) | (k << N2E_K_SHIFT);
            }
    }
    for (int i = 0; i < nn; ++i) T.nodes[no + i] = make_double4(X[i], Y[i], mv->z[i], mv->n[i]);
    MemberDev &D = T.mdev[mi];
    memset(&D, 0, sizeof(D));
    D.el_off = eo; D.ne = ne; D.ed_off = dof; D.ned = ned;
    D.eblk0 = eo / 256; D.eblks = pad256(ne) / 256;
    D.dblk0 = dof / 256; D.dblks = pad256(ned) / 256;
    D.tide_code = M.prm.tide_code;
    D.rain = M.prm.rain; D.cfl = M.prm.cfl; D.min_edge = M.min_edge; D.final_time = M.prm.final_time;
    D.tide_datum = M.prm.tide_datum; D.tide_amp = M.prm.tide_amp; D.tide_period = M.prm.tide_period;
    D.tide_phase = M.prm.tide_phase;
    D.sentinels = M.prm.bc_sentinels != 0;
    for (int b = 0; b < D.eblks; ++b) T.eblk_mem[D.eblk0 + b] = mi;
    for (int b = 0; b < D.dblks; ++b) T.dblk_mem[D.dblk0 + b] = mi;
    return DGR_OK;
}

static dgr_status create_members(int n, const dgr_mesh_view *mvs, const dgr_params *prms, dgr_handle **out)
{
    if (n <= 0 || !mvs || !prms || !out) return DGR_ERR_ARG;
    *out = nullptr;
    std::vector<MemberHost> mem(n);
    int pitch = 0, nn_tot = 0, ned_pad = 0, record_modify = 0;
    for (int i = 0; i < n; ++i) {
        const dgr_mesh_view *mv = &mvs[i];
        const dgr_params *p = &prms[i];
        const int ne = mv->num_el, nn = mv->num_nodes, ned = mv->num_edges;
        if (ne <= 0 || nn <= 0 || ned <= 0 || !mv->x || !mv->y || !mv->z || !mv->n || !mv->elto_vert ||
            !mv->edg_to_vert || !mv->edg_to_els || !mv->bdry)
            return DGR_ERR_ARG;
        // constants shared by every member
        if (p->g != prms[0].g || p->h0 != prms[0].h0 || p->velzero != prms[0].velzero ||
            (p->wetdry != 0) != (prms[0].wetdry != 0))
            return DGR_ERR_ARG;
        if ((long long)pitch + pad256(ne) > (1LL << 29) || (long long)ned_pad + pad256(ned) > (1LL << 29))
            return DGR_ERR_ARG;
        MemberHost &M = mem[i];
        M.ne = ne; M.nn = nn; M.ned = ned;
        M.el_off = pitch; M.nd_off = nn_tot; M.ed_off = ned_pad;
        M.prm = *p;
        // minEdgLength (extract_mesh_attributes.c:297-303)
        double mel = 9999999;
        for (int e = 0; e < ned; ++e) {
            const int n0 = mv->edg_to_vert[2 * e], n1 = mv->edg_to_vert[2 * e + 1];
            if (n0 < 0 || n0 >= nn || n1 < 0 || n1 >= nn) return DGR_ERR_MESH;
            const double dx = mv->x[n1] - mv->x[n0], dy = mv->y[n1] - mv->y[n0];
            const double len = std::sqrt(dx * dx + dy * dy);
            mel = (mel < len) ? mel : len;
        }
        M.min_edge = mel;
        pitch += pad256(ne);
        nn_tot += nn;
        ned_pad += pad256(ned);
        record_modify |= p->record_modify != 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DGR_ERR_NODEV;

    HostTables T;
    const size_t P = pitch;
    T.ev.assign(3 * P, 0); T.nb.assign(3 * P, 0); T.info.assign(3 * P, 0); T.eid.assign(3 * P, 0);
    T.off.assign((size_t)nn_tot + 1, 0);
    size_t n2e_tot = 0;
    for (int i = 0; i < n; ++i) n2e_tot += 3 * (size_t)mem[i].ne;
    T.n2e.assign(n2e_tot, 0);
    T.edges.assign(ned_pad, make_int4(0, 0, 0, 0));
    T.einfo.assign(ned_pad, 0);
    T.nodes.resize(nn_tot);
    T.bz.assign(ned_pad, 0.0); T.bq.assign(ned_pad, 0.0);
    T.mdev.resize(n);
    T.eblk_mem.assign(P / 256, 0);
    T.dblk_mem.assign((size_t)ned_pad / 256, 0);
    for (int i = 0; i < n; ++i) {
        dgr_status st = build_member(&mvs[i], mem[i], i, pitch, T);
        if (st) return st;
    }
    // per group of 64 nodes (a wave of the node kernels): the distinct
    // elements of its CSR entries, ascending, and each entry's position in
    // that list (three 10-bit positions per word)
    {
        const int ngrp = (nn_tot + 63) / 64;
        T.nl_off.assign((size_t)ngrp + 1, 0);
        T.loc_off.assign((size_t)nn_tot + 1, 0);
        std::vector<int> els;
        for (int g = 0; g < ngrp; ++g) {
            const int a = 64 * g, b = std::min(a + 64, nn_tot);
            els.clear();
            for (int j = T.off[a]; j < T.off[b]; ++j) els.push_back(T.n2e[j] & N2E_EL_MASK);
            std::sort(els.begin(), els.end());
            els.erase(std::unique(els.begin(), els.end()), els.end());
            T.nl_off[g] = (int)T.nl.size();
            T.nl.insert(T.nl.end(), els.begin(), els.end());
            for (int i = a; i < b; ++i) {
                T.loc_off[i] = (int)T.locw.size();
                unsigned word = 0;
                int t = 0;
                for (int j = T.off[i]; j < T.off[i + 1]; ++j) {
                    const int el = T.n2e[j] & N2E_EL_MASK;
                    const unsigned loc = (unsigned)(std::lower_bound(els.begin(), els.end(), el) - els.begin());
                    word |= (loc & 1023u) << (10 * t);
                    if (++t == 3) { T.locw.push_back(word); word = 0; t = 0; }
                }
                if (t) T.locw.push_back(word);
            }
        }
        T.nl_off[ngrp] = (int)T.nl.size();
        T.loc_off[nn_tot] = (int)T.locw.size();
        if (T.nl.empty()) T.nl.push_back(0);
        T.locw.push_back(0);   // padding: the node kernels read a node's second word unconditionally
    }

    dgr_handle *h = new dgr_handle();   // value-initialised: all pointers null
    h->nmem = n; h->pitch = pitch; h->nn = nn_tot; h->ned_pad = ned_pad;
    h->mem = mem;
    h->record_modify = record_modify;
    (void)hipGetDevice(&h->device);
#define DGR_TRY(call)                                          \
    do {                                                       \
        if ((call) != hipSuccess) { free_handle(h); return DGR_ERR_HIP; } \
    } while (0)
    DGR_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    hipStream_t s = h->stream;
    const int maxparts = (pitch > ned_pad ? pitch : ned_pad) / 256;
    DGR_TRY(dalloc(&h->d_nodes, nn_tot));
    DGR_TRY(dalloc(&h->d_zn, nn_tot));
    DGR_TRY(dalloc(&h->d_ev, 3 * P));
    DGR_TRY(dalloc(&h->d_nb, 3 * P));
    DGR_TRY(dalloc(&h->d_info, 3 * P));
    DGR_TRY(dalloc(&h->d_eid, 3 * P));
    DGR_TRY(dalloc(&h->d_n2e_off, (size_t)nn_tot + 1));
    DGR_TRY(dalloc(&h->d_n2e, n2e_tot));
    DGR_TRY(dalloc(&h->d_nl_off, T.nl_off.size()));
    DGR_TRY(dalloc(&h->d_nl, T.nl.size()));
    DGR_TRY(dalloc(&h->d_loc_off, T.loc_off.size()));
    DGR_TRY(dalloc(&h->d_locw, T.locw.size()));
    DGR_TRY(dalloc(&h->d_edges, ned_pad));
    DGR_TRY(dalloc(&h->d_einfo, ned_pad));
    DGR_TRY(dalloc(&h->d_egeo, ned_pad));
    DGR_TRY(dalloc(&h->d_mem, n));
    DGR_TRY(dalloc(&h->d_eblk_mem, P / 256));
    DGR_TRY(dalloc(&h->d_dblk_mem, (size_t)ned_pad / 256));
    DGR_TRY(dalloc(&h->d_bzeta, ned_pad));
    DGR_TRY(dalloc(&h->d_bqn, ned_pad));
    DGR_TRY(dalloc(&h->d_w, 9 * P));
    DGR_TRY(dalloc(&h->d_w1, 9 * P));
    DGR_TRY(dalloc(&h->d_means, 3 * P));
    DGR_TRY(dalloc(&h->d_nodemm, 6 * (size_t)nn_tot));
    DGR_TRY(dalloc(&h->d_fhat, 9 * P));
    DGR_TRY(dalloc(&h->d_wd, P));
    DGR_TRY(dalloc(&h->d_modify, 2 * P));
    DGR_TRY(dalloc(&h->d_ctrl, n));
    DGR_TRY(dalloc(&h->d_lamparts, 2 * (size_t)maxparts));
    DGR_TRY(dalloc(&h->d_nanparts, 2 * (size_t)maxparts));
    DGR_TRY(dalloc(&h->d_tiemax, (size_t)maxparts));
    DGR_TRY(dalloc(&h->d_tiecnt, (size_t)maxparts));
    DGR_TRY(dalloc(&h->d_wdn, P));
    DGR_TRY(hipMemcpyAsync(h->d_nodes, T.nodes.data(), nn_tot * sizeof(double4), hipMemcpyHostToDevice, s));
    std::vector<double> zn(nn_tot);
    for (int i = 0; i < nn_tot; ++i) zn[i] = T.nodes[i].z;
    DGR_TRY(hipMemcpyAsync(h->d_zn, zn.data(), nn_tot * sizeof(double), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_ev, T.ev.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_nb, T.nb.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_info, T.info.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_eid, T.eid.data(), 3 * P * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_n2e_off, T.off.data(), ((size_t)nn_tot + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_n2e, T.n2e.data(), n2e_tot * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_nl_off, T.nl_off.data(), T.nl_off.size() * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_nl, T.nl.data(), T.nl.size() * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_loc_off, T.loc_off.data(), T.loc_off.size() * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_locw, T.locw.data(), T.locw.size() * sizeof(unsigned), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_edges, T.edges.data(), ned_pad * sizeof(int4), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_einfo, T.einfo.data(), ned_pad * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_mem, T.mdev.data(), n * sizeof(MemberDev), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_eblk_mem, T.eblk_mem.data(), P / 256 * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_dblk_mem, T.dblk_mem.data(), (size_t)ned_pad / 256 * sizeof(int),
                           hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_bzeta, T.bz.data(), ned_pad * sizeof(double), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemcpyAsync(h->d_bqn, T.bq.data(), ned_pad * sizeof(double), hipMemcpyHostToDevice, s));
    DGR_TRY(hipMemsetAsync(h->d_w, 0, 9 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_w1, 0, 9 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_means, 0, 3 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_fhat, 0, 9 * P * sizeof(double), s));
    DGR_TRY(hipMemsetAsync(h->d_modify, 0, 2 * P, s));
    DGR_TRY(hipMemsetAsync(h->d_lamparts, 0, 2 * (size_t)maxparts * sizeof(unsigned long long), s));
    DGR_TRY(hipMemsetAsync(h->d_nanparts, 0xff, 2 * (size_t)maxparts * sizeof(int), s));
    {
        std::vector<int> wd(P, -1);   // initialize.c:126
        DGR_TRY(hipMemcpyAsync(h->d_wd, wd.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
        DGR_TRY(hipMemcpyAsync(h->d_wdn, wd.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
        std::vector<Ctrl> c(n);
        for (int i = 0; i < n; ++i) {
            memset(&c[i], 0, sizeof(Ctrl));
            c[i].t = 0.0;
            c[i].dt = prms[i].dt0;
            c[i].err_key = ~0ULL;
            c[i].nan_pos[0] = c[i].nan_pos[1] = -1;
        }
        DGR_TRY(hipMemcpyAsync(h->d_ctrl, c.data(), n * sizeof(Ctrl), hipMemcpyHostToDevice, s));
        DGR_TRY(hipStreamSynchronize(s));
    }
#undef DGR_TRY
    MeshDev &m = h->m;
    m.pitch = pitch; m.nn = nn_tot; m.ned_pad = ned_pad; m.nmem = n;
    m.nodes = h->d_nodes; m.zn = h->d_zn; m.ev = h->d_ev; m.nb = h->d_nb; m.info = h->d_info; m.eid = h->d_eid;
    m.n2e_off = h->d_n2e_off; m.n2e = h->d_n2e; m.edges = h->d_edges;
    m.nl_off = h->d_nl_off; m.nl = h->d_nl; m.loc_off = h->d_loc_off; m.locw = h->d_locw; m.einfo = h->d_einfo;
    m.egeo = h->d_egeo;
    m.mem = h->d_mem; m.eblk_mem = h->d_eblk_mem; m.dblk_mem = h->d_dblk_mem;
    m.wg = 256;
    PhysDev &ph = h->ph;
    ph.g = prms[0].g; ph.h0 = prms[0].h0; ph.velzero = prms[0].velzero; ph.wetdry = prms[0].wetdry;
    ph.lamparts = h->d_lamparts; ph.nanparts = h->d_nanparts; ph.maxparts = maxparts;
    ph.tiemax = h->d_tiemax; ph.tiecnt = h->d_tiecnt;
    size_t maxne = 0;
    for (int i = 0; i < n; ++i) maxne = maxne > (size_t)mem[i].ne ? maxne : (size_t)mem[i].ne;
    h->h_planes.resize(9 * maxne);
    launch_edge_geometry(h->m, h->d_egeo, s);
    // tide value for t = 0
    launch_step_begin(h->m, h->d_ctrl, s);
    if (hipStreamSynchronize(s) != hipSuccess) { free_handle(h); return DGR_ERR_HIP; }
    *out = h;
    return DGR_OK;
}

extern "C" dgr_status dgr_create(const dgr_mesh_view *mesh, const dgr_params *params, dgr_handle **out)
{
    return create_members(1, mesh, params, out);
}

extern "C" dgr_status dgr_ensemble_create(int n, const dgr_mesh_view *meshes, const dgr_params *params,
                                          dgr_handle **out)
{
    return create_members(n, meshes, params, out);
}

extern "C" dgr_status dgr_destroy(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    (void)hipStreamSynchronize(h->stream);
    return free_handle(h);
}

extern "C" dgr_status dgr_num_members(dgr_handle *h, int *n)
{
    if (!h || !n) return DGR_ERR_ARG;
    *n = h->nmem;
    return DGR_OK;
}

static inline bool bad_member(dgr_handle *h, int mi) { return !h || mi < 0 || mi >= h->nmem; }

// reference layout (el*3+k) <-> device planes of member mi
static dgr_status planes_up(dgr_handle *h, int mi, double *dst, const double *a, const double *b, const double *c)
{
    const MemberHost &M = h->mem[mi];
    const size_t ne = M.ne, P = h->pitch;
    const double *src[3] = {a, b, c};
    for (int f = 0; f < 3; ++f)
        for (size_t el = 0; el < ne; ++el)
            for (int k = 0; k < 3; ++k) h->h_planes[(3 * f + k) * ne + el] = src[f][3 * el + k];
    for (int pl = 0; pl < 9; ++pl)
        DGR_HIP(hipMemcpyAsync(dst + pl * P + M.el_off, h->h_planes.data() + pl * ne, ne * sizeof(double),
                               hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

static dgr_status planes_down(dgr_handle *h, int mi, const double *srcd, double *a, double *b, double *c)
{
    const MemberHost &M = h->mem[mi];
    const size_t ne = M.ne, P = h->pitch;
    for (int pl = 0; pl < 9; ++pl)
        DGR_HIP(hipMemcpyAsync(h->h_planes.data() + pl * ne, srcd + pl * P + M.el_off, ne * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    double *dst[3] = {a, b, c};
    for (int f = 0; f < 3; ++f)
        if (dst[f])
            for (size_t el = 0; el < ne; ++el)
                for (int k = 0; k < 3; ++k) dst[f][3 * el + k] = h->h_planes[(3 * f + k) * ne + el];
    return DGR_OK;
}

extern "C" dgr_status dgr_member_set_state(dgr_handle *h, int mi, const double *z, const double *qx, const double *qy)
{
    if (bad_member(h, mi) || !z || !qx || !qy) return DGR_ERR_ARG;
    const unsigned long long none = ~0ULL;   // a new state clears a recorded error
    DGR_HIP(hipMemcpyAsync(&h->d_ctrl[mi].err_key, &none, sizeof(none), hipMemcpyHostToDevice, h->stream));
    return planes_up(h, mi, h->d_w, z, qx, qy);
}

extern "C" dgr_status dgr_member_get_state(dgr_handle *h, int mi, double *z, double *qx, double *qy)
{
    if (bad_member(h, mi)) return DGR_ERR_ARG;
    return planes_down(h, mi, h->d_w, z, qx, qy);
}

extern "C" dgr_status dgr_member_set_wd(dgr_handle *h, int mi, const int *wd)
{
    if (bad_member(h, mi) || !wd) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    DGR_HIP(hipMemcpyAsync(h->d_wd + M.el_off, wd, M.ne * sizeof(int), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_member_get_wd(dgr_handle *h, int mi, int *wd)
{
    if (bad_member(h, mi) || !wd) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    DGR_HIP(hipMemcpyAsync(wd, h->d_wd + M.el_off, M.ne * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_member_get_modify(dgr_handle *h, int mi, int stage, unsigned char *mod)
{
    if (bad_member(h, mi) || !mod || stage < 0 || stage > 1) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    DGR_HIP(hipMemcpyAsync(mod, h->d_modify + (size_t)stage * h->pitch + M.el_off, M.ne, hipMemcpyDeviceToHost,
                           h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

// Hydrographs: every member's tables are concatenated into one device array
// (MeshDev::hydro), rebuilt on each change; the captured step graph holds the
// mesh descriptor by value, so it is dropped.
static dgr_status refresh_forcing(dgr_handle *h, int mi);

extern "C" dgr_status dgr_member_set_hydrograph(dgr_handle *h, int mi, int code, int npts, const double *t,
                                                const double *zeta_b, const double *qn_b)
{
    if (bad_member(h, mi) || code < 1 || code > 3 || npts < 0 || (npts > 0 && (!t || !zeta_b || !qn_b)))
        return DGR_ERR_ARG;
    if (code == h->mem[mi].prm.tide_code) return DGR_ERR_ARG;
    for (int i = 1; i < npts; ++i)
        if (!(t[i] > t[i - 1])) return DGR_ERR_ARG;
    std::vector<double> &v = h->mem[mi].hyd[code - 1];
    v.assign(3 * (size_t)npts, 0.0);
    for (int i = 0; i < npts; ++i) { v[i] = t[i]; v[npts + i] = zeta_b[i]; v[2 * npts + i] = qn_b[i]; }
    std::vector<MemberDev> md(h->nmem);
    DGR_HIP(hipMemcpyAsync(md.data(), h->d_mem, h->nmem * sizeof(MemberDev), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    std::vector<double> all;
    for (int i = 0; i < h->nmem; ++i)
        for (int c = 0; c < 3; ++c) {
            const std::vector<double> &hv = h->mem[i].hyd[c];
            md[i].hyd_n[c] = (int)(hv.size() / 3);
            md[i].hyd_off[c] = (int)all.size();
            all.insert(all.end(), hv.begin(), hv.end());
        }
    if (h->d_hydro) {
        DGR_HIP(hipStreamSynchronize(h->stream));
        DGR_HIP(hipFree(h->d_hydro));
        h->d_hydro = nullptr;
    }
    if (!all.empty()) {
        DGR_HIP(dalloc(&h->d_hydro, all.size()));
        DGR_HIP(hipMemcpyAsync(h->d_hydro, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
    }
    DGR_HIP(hipMemcpyAsync(h->d_mem, md.data(), h->nmem * sizeof(MemberDev), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    h->m.hydro = h->d_hydro;
    if (h->graph_exec) {
        DGR_HIP(hipGraphExecDestroy(h->graph_exec));
        h->graph_exec = nullptr;
    }
    return refresh_forcing(h, mi);
}

extern "C" dgr_status dgr_set_hydrograph(dgr_handle *h, int code, int npts, const double *t, const double *zeta_b,
                                         const double *qn_b)
{
    return dgr_member_set_hydrograph(h, 0, code, npts, t, zeta_b, qn_b);
}

extern "C" dgr_status dgr_member_set_infiltration(dgr_handle *h, int mi, double K, double psi, double theta_e,
                                                  double s_e)
{
    if (bad_member(h, mi)) return DGR_ERR_ARG;
    double kk = 0.0, pp = 0.0;
    if (K > 0.0) {
        if (!(psi > 0.0) || !(theta_e > 0.0) || theta_e > 1.0 || s_e < 0.0 || !(s_e < 1.0) || K > 1e300 ||
            psi > 1e300)
            return DGR_ERR_ARG;
        kk = K;
        pp = psi * ((1 - s_e) * theta_e);   // KinField/Infiltration.c:16,49
    }
    MemberHost &M = h->mem[mi];
    M.infil_K = kk;
    M.infil_P = pp;
    // the members' device records are read at every step; no graph holds them
    MemberDev md;
    DGR_HIP(hipMemcpyAsync(&md, h->d_mem + mi, sizeof(MemberDev), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    md.infil_K = kk;
    md.infil_P = pp;
    DGR_HIP(hipMemcpyAsync(h->d_mem + mi, &md, sizeof(MemberDev), hipMemcpyHostToDevice, h->stream));
    return refresh_forcing(h, mi);
}

extern "C" dgr_status dgr_set_infiltration(dgr_handle *h, double K, double psi, double theta_e, double s_e)
{
    return dgr_member_set_infiltration(h, 0, K, psi, theta_e, s_e);
}

extern "C" dgr_status dgr_member_set_boundary(dgr_handle *h, int mi, const double *bz, const double *bq)
{
    if (bad_member(h, mi) || !bz || !bq) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    DGR_HIP(hipMemcpyAsync(h->d_bzeta + M.ed_off, bz, M.ned * sizeof(double), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(h->d_bqn + M.ed_off, bq, M.ned * sizeof(double), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_member_get_boundary(dgr_handle *h, int mi, double *bz, double *bq)
{
    if (bad_member(h, mi) || !bz || !bq) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    DGR_HIP(hipMemcpyAsync(bz, h->d_bzeta + M.ed_off, M.ned * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipMemcpyAsync(bq, h->d_bqn + M.ed_off, M.ned * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_get_boundary(dgr_handle *h, double *bz, double *bq)
{
    return dgr_member_get_boundary(h, 0, bz, bq);
}

extern "C" dgr_status dgr_couple_channel_ends(dgr_handle *h, int nends, const int *end_off, const int *edge_member,
                                              const int *edge_local, const int *end_dir, const double *end_A,
                                              const double *end_Q, double *ghost_A, double *ghost_Q)
{
    if (!h || nends < 0) return DGR_ERR_ARG;
    if (nends == 0) return DGR_OK;
    if (!end_off || !end_dir || !end_A || !end_Q || !ghost_A || !ghost_Q || end_off[0] != 0) return DGR_ERR_ARG;
    for (int e = 0; e < nends; ++e)
        if (end_off[e + 1] < end_off[e] || (end_dir[e] != 1 && end_dir[e] != -1)) return DGR_ERR_ARG;
    const int ne = end_off[nends];
    if (ne > 0 && (!edge_member || !edge_local)) return DGR_ERR_ARG;
    for (int j = 0; j < ne; ++j)
        if (bad_member(h, edge_member[j]) || edge_local[j] < 0 || edge_local[j] >= h->mem[edge_member[j]].ned)
            return DGR_ERR_ARG;
    // spec: end_off, members, local edges, directions; vals: A, Q, then the
    // ghost values come back in the same buffer
    const size_t ns = (size_t)nends + 1 + 2 * (size_t)ne + nends, nv = 4 * (size_t)nends;
    if (h->cpl_spec_n < ns) {
        if (h->d_cpl_spec) DGR_HIP(hipFree(h->d_cpl_spec));
        h->d_cpl_spec = nullptr;
        h->cpl_spec_n = 0;
        DGR_HIP(dalloc(&h->d_cpl_spec, ns));
        h->cpl_spec_n = ns;
    }
    if (h->cpl_vals_n < nv) {
        if (h->d_cpl_vals) DGR_HIP(hipFree(h->d_cpl_vals));
        h->d_cpl_vals = nullptr;
        h->cpl_vals_n = 0;
        DGR_HIP(dalloc(&h->d_cpl_vals, nv));
        h->cpl_vals_n = nv;
    }
    std::vector<int> spec(ns);
    std::vector<double> vals(nv);
    std::copy(end_off, end_off + nends + 1, spec.begin());
    std::copy(edge_member, edge_member + ne, spec.begin() + nends + 1);
    std::copy(edge_local, edge_local + ne, spec.begin() + nends + 1 + ne);
    std::copy(end_dir, end_dir + nends, spec.begin() + nends + 1 + 2 * ne);
    std::copy(end_A, end_A + nends, vals.begin());
    std::copy(end_Q, end_Q + nends, vals.begin() + nends);
    hipStream_t s = h->stream;
    DGR_HIP(hipMemcpyAsync(h->d_cpl_spec, spec.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_HIP(hipMemcpyAsync(h->d_cpl_vals, vals.data(), 2 * nends * sizeof(double), hipMemcpyHostToDevice, s));
    double *ghost = h->d_cpl_vals + 2 * (size_t)nends;
    launch_couple_ends(h->m, h->d_w, h->d_bzeta, h->d_bqn, nends, h->d_cpl_spec, h->d_cpl_vals, ghost, s);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipMemcpyAsync(ghost_A, ghost, nends * sizeof(double), hipMemcpyDeviceToHost, s));
    DGR_HIP(hipMemcpyAsync(ghost_Q, ghost + nends, nends * sizeof(double), hipMemcpyDeviceToHost, s));
    DGR_HIP(hipStreamSynchronize(s));
    return DGR_OK;
}

extern "C" dgr_status dgr_couple_dgshed(dgr_handle *h, int n, const int *edge_member, const int *edge_local,
                                       const int *conn, const double *chan, const double *geom, double *swe,
                                       double *ghost)
{
    if (!h || n < 0) return DGR_ERR_ARG;
    if (n == 0) return DGR_OK;
    if (!edge_member || !edge_local || !conn || !chan || !geom || !ghost) return DGR_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (bad_member(h, edge_member[i]) || edge_local[i] < 0 || edge_local[i] >= h->mem[edge_member[i]].ned ||
            (conn[i] != 0 && conn[i] != 1))
            return DGR_ERR_ARG;
    // spec: members, local edges, conn, flags; vals: chan, geom, swe, ghost
    const size_t ns = 4 * (size_t)n, nv = 14 * (size_t)n;
    if (h->cpl_spec_n < ns) {
        if (h->d_cpl_spec) DGR_HIP(hipFree(h->d_cpl_spec));
        h->d_cpl_spec = nullptr;
        h->cpl_spec_n = 0;
        DGR_HIP(dalloc(&h->d_cpl_spec, ns));
        h->cpl_spec_n = ns;
    }
    if (h->cpl_vals_n < nv) {
        if (h->d_cpl_vals) DGR_HIP(hipFree(h->d_cpl_vals));
        h->d_cpl_vals = nullptr;
        h->cpl_vals_n = 0;
        DGR_HIP(dalloc(&h->d_cpl_vals, nv));
        h->cpl_vals_n = nv;
    }
    std::vector<int> spec(ns, 0);
    std::copy(edge_member, edge_member + n, spec.begin());
    std::copy(edge_local, edge_local + n, spec.begin() + n);
    std::copy(conn, conn + n, spec.begin() + 2 * n);
    std::vector<double> vals(10 * (size_t)n);
    std::copy(chan, chan + 6 * (size_t)n, vals.begin());
    std::copy(geom, geom + 4 * (size_t)n, vals.begin() + 6 * (size_t)n);
    hipStream_t s = h->stream;
    DGR_HIP(hipMemcpyAsync(h->d_cpl_spec, spec.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    DGR_HIP(hipMemcpyAsync(h->d_cpl_vals, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_couple_dgshed(h->m, h->ph, h->d_w, h->d_bzeta, h->d_bqn, n, h->d_cpl_spec, h->d_cpl_vals, s);
    DGR_HIP(hipGetLastError());
    std::vector<double> out(4 * (size_t)n);
    DGR_HIP(hipMemcpyAsync(out.data(), h->d_cpl_vals + 10 * (size_t)n, out.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    DGR_HIP(hipMemcpyAsync(spec.data() + 3 * n, h->d_cpl_spec + 3 * (size_t)n, n * sizeof(int), hipMemcpyDeviceToHost,
                           s));
    DGR_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i)
        if (spec[3 * (size_t)n + i]) return DGR_ERR_ARG;   // not a boundary edge (left unchanged)
    if (swe) std::copy(out.begin(), out.begin() + 2 * (size_t)n, swe);
    std::copy(out.begin() + 2 * (size_t)n, out.end(), ghost);
    return DGR_OK;
}

dgr_status read_ctrl(dgr_handle *h, int mi, Ctrl *c)
{
    DGR_HIP(hipMemcpyAsync(c, h->d_ctrl + mi, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

// Member mi's time-dependent forcing (tide, hydrograph values, infiltration
// rate) at its current t and dt, as k_step_begin evaluates it, so that single
// operators see the forcing of the handle's time; the other members' Ctrl and
// every step counter are left as they were.
static dgr_status refresh_forcing(dgr_handle *h, int mi)
{
    std::vector<Ctrl> all(h->nmem), after(h->nmem);
    DGR_HIP(hipMemcpyAsync(all.data(), h->d_ctrl, h->nmem * sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    launch_step_begin(h->m, h->d_ctrl, h->stream);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipMemcpyAsync(after.data(), h->d_ctrl, h->nmem * sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    all[mi].tide = after[mi].tide;
    all[mi].infil = after[mi].infil;
    for (int k = 0; k < 3; ++k) {
        all[mi].hyd_z[k] = after[mi].hyd_z[k];
        all[mi].hyd_q[k] = after[mi].hyd_q[k];
    }
    DGR_HIP(hipMemcpyAsync(h->d_ctrl, all.data(), h->nmem * sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_member_set_time(dgr_handle *h, int mi, double t, double dt)
{
    if (bad_member(h, mi)) return DGR_ERR_ARG;
    // t, dt, then the forcing at t; the step counter (which only orders
    // error keys) is left as it was
    Ctrl c;
    dgr_status st = read_ctrl(h, mi, &c);
    if (st) return st;
    c.t = t;
    c.dt = dt;
    DGR_HIP(hipMemcpyAsync(h->d_ctrl + mi, &c, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
    return refresh_forcing(h, mi);
}

static void fill_report(const Ctrl &c, dgr_report *rep)
{
    rep->t = c.t;
    rep->dt = c.dt;
    rep->max_lambda = maxlam_of(c, 1);
    rep->error_code = err_code_of(c.err_key);
    rep->error_where = err_where_of(c.err_key);
}

extern "C" dgr_status dgr_member_report(dgr_handle *h, int mi, dgr_report *rep)
{
    if (bad_member(h, mi) || !rep) return DGR_ERR_ARG;
    Ctrl c;
    dgr_status st = read_ctrl(h, mi, &c);
    if (st) return st;
    fill_report(c, rep);
    return rep->error_code ? DGR_ERR_NAN : DGR_OK;
}

// any member with a recorded error
dgr_status any_error(dgr_handle *h)
{
    std::vector<Ctrl> c(h->nmem);
    DGR_HIP(hipMemcpyAsync(c.data(), h->d_ctrl, h->nmem * sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    for (const Ctrl &x : c)
        if (err_code_of(x.err_key)) return DGR_ERR_NAN;
    return DGR_OK;
}

// single-instance accessors: member 0
extern "C" dgr_status dgr_set_state(dgr_handle *h, const double *z, const double *qx, const double *qy)
{
    return dgr_member_set_state(h, 0, z, qx, qy);
}
extern "C" dgr_status dgr_get_state(dgr_handle *h, double *z, double *qx, double *qy)
{
    return dgr_member_get_state(h, 0, z, qx, qy);
}
extern "C" dgr_status dgr_set_wd(dgr_handle *h, const int *wd) { return dgr_member_set_wd(h, 0, wd); }
extern "C" dgr_status dgr_get_wd(dgr_handle *h, int *wd) { return dgr_member_get_wd(h, 0, wd); }
extern "C" dgr_status dgr_get_modify(dgr_handle *h, int stage, unsigned char *mod)
{
    return dgr_member_get_modify(h, 0, stage, mod);
}
extern "C" dgr_status dgr_set_boundary(dgr_handle *h, const double *bz, const double *bq)
{
    return dgr_member_set_boundary(h, 0, bz, bq);
}
extern "C" dgr_status dgr_set_time(dgr_handle *h, double t, double dt) { return dgr_member_set_time(h, 0, t, dt); }

unsigned char *modify_ptr(dgr_handle *h, int stage)
{
    return h->record_modify ? h->d_modify + (size_t)stage * h->pitch : nullptr;
}

extern "C" dgr_status dgr_compute_rhs(dgr_handle *h, double *rz, double *rqx, double *rqy, double *f1, double *f2,
                                      double *f3, double *max_lambda)
{
    if (!h) return DGR_ERR_ARG;
    if (!h->d_dbg) DGR_HIP(dalloc(&h->d_dbg, 18 * (size_t)h->pitch));
    const bool wdon = h->ph.wetdry != 0;
    launch_reset_maxlam(h->m, h->d_ctrl, h->stream);
    h->kind[0] = launch_stage(h->variant, 0, wdon, h->m, h->ph, h->d_w, nullptr, nullptr, h->d_bzeta, h->d_bqn,
                              h->d_wd, nullptr, h->d_fhat, h->d_ctrl, 0, h->d_dbg, h->d_dbg + 9 * (size_t)h->pitch,
                              h->stream);
    // element partials: repair the fold after a NaN operand before reducing
    // (the edge kernel's partials carry the NaN position themselves)
    if (h->kind[0] == PARTS_ELEM)
        launch_stage(h->variant, 3, wdon, h->m, h->ph, h->d_w, nullptr, nullptr, h->d_bzeta, h->d_bqn, h->d_wd,
                     nullptr, h->d_fhat, h->d_ctrl, 0, nullptr, nullptr, h->stream);
    launch_lam_reduce(h->m, h->ph, h->d_ctrl, 0, h->kind[0], h->stream);
    DGR_HIP(hipGetLastError());
    if (rz || rqx || rqy) {
        dgr_status st = planes_down(h, 0, h->d_dbg, rz, rqx, rqy);
        if (st) return st;
    }
    if (f1 || f2 || f3) {
        dgr_status st = planes_down(h, 0, h->d_dbg + 9 * (size_t)h->pitch, f1, f2, f3);
        if (st) return st;
    }
    Ctrl c;
    dgr_status st = read_ctrl(h, 0, &c);
    if (st) return st;
    if (max_lambda) *max_lambda = maxlam_of(c, 0);
    return any_error(h);
}

extern "C" dgr_status dgr_limit(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_means(h->m, h->d_w, h->d_means, h->stream);
    launch_node_minmax(h->m, h->d_means, h->d_nodemm, nullptr, h->stream);
    launch_limit(false, h->m, h->ph, h->d_w, h->d_w, h->d_nodemm, modify_ptr(h, 0), h->d_ctrl, 0, nullptr, nullptr,
                 h->stream);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_wetdry_status(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_wd_status(h->m, h->ph, h->d_w, h->d_wd, h->d_wd, h->stream);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_pdop(dgr_handle *h)
{
    if (!h) return DGR_ERR_ARG;
    launch_pdop(h->m, h->ph, h->d_w, h->d_ctrl, h->stream);
    DGR_HIP(hipGetLastError());
    return any_error(h);
}

// event pairs around the flux/RHS/update launches when profiling is on
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
// for every member at once.  The wetDryStatus2D call between the stages runs
// inside the stage-1 limiter kernel.  b == nullptr: the plain in-place buffers.
void enqueue_stages(dgr_handle *h, const StepBufs *b, bool fuse_end, StepRec *rec_end, StepRec *rec_next)
{
    const bool wdon = h->ph.wetdry != 0;
    hipStream_t s = h->stream;
    const StepBufs plain = {h->d_w, h->d_w1, h->d_w1, h->d_w, h->d_w, h->d_wd, h->d_wd, h->d_wd, nullptr, false,
                            nullptr, nullptr, nullptr, nullptr, false};
    if (!b) b = &plain;
    if (wdon && !b->wda_ready) launch_wd_status(h->m, h->ph, b->w0, b->wdprev, b->wda, s);
    mark(h);
    h->kind[0] = launch_stage(h->variant, 1, wdon, h->m, h->ph, b->w0, b->u1, nullptr, h->d_bzeta, h->d_bqn, b->wda,
                              h->d_means, h->d_fhat, h->d_ctrl, 0, nullptr, nullptr, s);
    mark(h);
    double *nm1 = b->nodemm1 ? b->nodemm1 : h->d_nodemm;
    launch_node_minmax(h->m, h->d_means, nm1, b->nodearg1, s);
    launch_limit(wdon, h->m, h->ph, b->u1, b->p1, nm1, modify_ptr(h, 0), h->d_ctrl, 0, b->wda,
                 wdon ? b->wdb : nullptr, s);
    const int *wd2 = wdon ? b->wdb : b->wda;
    mark(h);
    // a taped step's edge kernel also counts the stage-2 max-|lambda| ties
    // for the reverse (k_step_end records them)
    const bool ties = b->taped;
    h->kind[1] = launch_stage(h->variant, 2, wdon, h->m, h->ph, b->p1, b->u2, b->w0, h->d_bzeta, h->d_bqn, wd2,
                              h->d_means, h->d_fhat, h->d_ctrl, 1, nullptr, nullptr, s, ties);
    h->ties_last = ties && h->kind[1] == PARTS_EDGE;
    mark(h);
    // stage 2's max |lambda| sets the next dt.  Element partials: repair the
    // fold if a NaN operand was seen (edge partials carry the NaN position).
    if (h->kind[1] == PARTS_ELEM)
        launch_stage(h->variant, 3, wdon, h->m, h->ph, b->p1, nullptr, nullptr, h->d_bzeta, h->d_bqn, wd2, nullptr,
                     h->d_fhat, h->d_ctrl, 1, nullptr, nullptr, s);
    double *nm2 = b->nodemm2 ? b->nodemm2 : h->d_nodemm;
    if (fuse_end)
        launch_node_minmax_end(h->m, h->ph, h->d_means, nm2, b->nodearg2, h->d_ctrl, h->kind[1], rec_end, rec_next,
                               h->ties_last, s);
    else
        launch_node_minmax(h->m, h->d_means, nm2, b->nodearg2, s);
    const bool wdn = wdon && b->wdnext;
    launch_limit(wdon, h->m, h->ph, b->u2, b->wout, nm2, modify_ptr(h, 1), h->d_ctrl, 1,
                 wdn ? wd2 : nullptr, wdn ? b->wdnext : nullptr, s, fuse_end ? 1 : 0);
}

void enqueue_step(dgr_handle *h, const StepBufs *b)
{
    launch_step_begin(h->m, h->d_ctrl, h->stream);
    enqueue_stages(h, b);
    launch_step_end(h->m, h->ph, h->d_ctrl, h->kind[1], false, h->stream);
}

// The steps of dgr_run.  Before the first: the step bookkeeping and the
// start flags into d_wdn.  Each step then takes its start flags from d_wdn,
// keeps the stage-2 flags in d_wd (as the reference leaves WD after a step),
// writes the next step's start flags into d_wdn from its stage-2 limiter and
// ends with the next step's bookkeeping: no separate status or step-begin
// launch per step.
void enqueue_run_prologue(dgr_handle *h)
{
    launch_step_begin(h->m, h->d_ctrl, h->stream);
    if (h->ph.wetdry) launch_wd_status(h->m, h->ph, h->d_w, h->d_wd, h->d_wdn, h->stream);
}

// a run step of the persistent variant: two cooperative launches, the
// buffers of the plain run step (see enqueue_run_step)
static void enqueue_run_step_persist(dgr_handle *h)
{
    const bool wdon = h->ph.wetdry != 0;
    PersistArgs a{};
    a.bzeta = h->d_bzeta;
    a.bqn = h->d_bqn;
    a.fhat = h->d_fhat;
    a.means = h->d_means;
    a.nodemm = h->d_nodemm;
    a.ctrls = h->d_ctrl;
    // stage 1: w -> w1 (flags d_wdn; the limiter writes the mid-step flags to d_wd)
    a.win = h->d_w;
    a.u = h->d_w1;
    a.wbase = nullptr;
    a.wd = h->d_wdn;
    a.lout = h->d_w1;
    a.modify = modify_ptr(h, 0);
    a.wd_lin = h->d_wdn;
    a.wd_lout = wdon ? h->d_wd : nullptr;
    a.next = 0;
    hipError_t e = launch_stage_persist(1, wdon, h->m, h->ph, a, h->stream);
    // stage 2: w1 -> w in place (flags d_wd; the limiter writes the next
    // step's start flags to d_wdn), then the step end
    a.win = h->d_w1;
    a.u = h->d_w;
    a.wbase = h->d_w;
    a.wd = wdon ? h->d_wd : h->d_wdn;
    a.lout = h->d_w;
    a.modify = modify_ptr(h, 1);
    a.wd_lin = wdon ? h->d_wd : nullptr;
    a.wd_lout = wdon ? h->d_wdn : nullptr;
    a.next = 1;
    if (e == hipSuccess) e = launch_stage_persist(2, wdon, h->m, h->ph, a, h->stream);
    if (e != hipSuccess) fprintf(stderr, "libdgrain: persistent stage launch failed: %s\n", hipGetErrorString(e));
    h->kind[0] = h->kind[1] = PARTS_EDGE;
    h->ties_last = false;
}

void enqueue_run_step(dgr_handle *h, const StepBufs *b, StepRec *rec_end, StepRec *rec_next)
{
    if (!b && !rec_end && !rec_next && h->variant == DGR_VARIANT_PERSIST) {
        enqueue_run_step_persist(h);
        return;
    }
    const StepBufs run = {h->d_w, h->d_w1, h->d_w1, h->d_w, h->d_w, h->d_wd, h->d_wdn, h->d_wd, h->d_wdn, true,
                          nullptr, nullptr, nullptr, nullptr, false};
    // the step ends inside its stage-2 node bounds launch
    enqueue_stages(h, b ? b : &run, true, rec_end, rec_next);
}

extern "C" dgr_status dgr_one_time_step(dgr_handle *h, double t, double dt)
{
    if (!h) return DGR_ERR_ARG;
    (void)t;
    launch_step_mark(h->m, h->d_ctrl, dt, h->stream);
    enqueue_stages(h, nullptr);
    launch_lam_reduce(h->m, h->ph, h->d_ctrl, 1, h->kind[1], h->stream);
    DGR_HIP(hipGetLastError());
    return any_error(h);
}

// one step of every member captured as a hipGraph (rebuilt when the variant changes)
static dgr_status step_graph(dgr_handle *h)
{
    if (h->graph_exec && h->graph_variant == h->variant && h->graph_wg == h->m.wg) return DGR_OK;
    if (h->graph_exec) {
        DGR_HIP(hipGraphExecDestroy(h->graph_exec));
        h->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    DGR_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_run_step(h);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        fprintf(stderr, "libdgrain: step capture failed: %s\n", hipGetErrorString(le != hipSuccess ? le : ce));
        return DGR_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        h->graph_exec = nullptr;
        fprintf(stderr, "libdgrain: graph instantiate failed: %s\n", hipGetErrorString(ie));
        return DGR_ERR_HIP;
    }
    h->graph_variant = h->variant;
    h->graph_wg = h->m.wg;
    return DGR_OK;
}

extern "C" dgr_status dgr_set_graph(dgr_handle *h, int on)
{
    if (!h) return DGR_ERR_ARG;
    h->use_graph = on != 0;
    return DGR_OK;
}

extern "C" dgr_status dgr_run(dgr_handle *h, int nsteps, dgr_report *rep)
{
    if (!h || nsteps < 0) return DGR_ERR_ARG;
    if (nsteps > 0 && h->use_graph && !h->profiling && h->variant != DGR_VARIANT_PERSIST) {
        dgr_status st = step_graph(h);
        if (st) return st;
        enqueue_run_prologue(h);
        for (int n = 0; n < nsteps; ++n) DGR_HIP(hipGraphLaunch(h->graph_exec, h->stream));
    } else if (nsteps > 0) {
        enqueue_run_prologue(h);
        for (int n = 0; n < nsteps; ++n) enqueue_run_step(h);
    }
    DGR_HIP(hipGetLastError());
    if (rep) {
        Ctrl c;
        dgr_status st = read_ctrl(h, 0, &c);
        if (st) return st;
        fill_report(c, rep);
    }
    return any_error(h);
}

extern "C" dgr_status dgr_get_report(dgr_handle *h, dgr_report *rep) { return dgr_member_report(h, 0, rep); }

extern "C" dgr_status dgr_set_final_time(dgr_handle *h, double final_time)
{
    if (!h) return DGR_ERR_ARG;
    std::vector<MemberDev> md(h->nmem);
    DGR_HIP(hipMemcpyAsync(md.data(), h->d_mem, h->nmem * sizeof(MemberDev), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->nmem; ++i) {
        h->mem[i].prm.final_time = final_time;
        md[i].final_time = final_time;
    }
    DGR_HIP(hipMemcpyAsync(h->d_mem, md.data(), h->nmem * sizeof(MemberDev), hipMemcpyHostToDevice, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_last_error(dgr_handle *h, int *code, int *where)
{
    if (!h) return DGR_ERR_ARG;
    Ctrl c;
    dgr_status st = read_ctrl(h, 0, &c);
    if (st) return st;
    if (code) *code = err_code_of(c.err_key);
    if (where) *where = err_where_of(c.err_key);
    return DGR_OK;
}

extern "C" dgr_status dgr_get_stream(dgr_handle *h, void **stream)
{
    if (!h || !stream) return DGR_ERR_ARG;
    *stream = (void *)h->stream;
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

// Restart file of one member: "DGRCKPT1", int32 num_el, double t, dt, then
// zeta, Qx, Qy (3*num_el doubles each, reference layout) and WD (num_el
// int32), host byte order.
extern "C" dgr_status dgr_checkpoint_write(dgr_handle *h, int mi, const char *path)
{
    if (bad_member(h, mi) || !path) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    const size_t n3 = 3 * (size_t)M.ne;
    std::vector<double> st(3 * n3);
    std::vector<int> wd(M.ne);
    dgr_status r = dgr_member_get_state(h, mi, st.data(), st.data() + n3, st.data() + 2 * n3);
    if (r) return r;
    r = dgr_member_get_wd(h, mi, wd.data());
    if (r) return r;
    Ctrl c;
    r = read_ctrl(h, mi, &c);
    if (r) return r;
    FILE *f = fopen(path, "wb");
    if (!f) return DGR_ERR_ARG;
    const int ne = M.ne;
    bool ok = fwrite("DGRCKPT1", 1, 8, f) == 8 && fwrite(&ne, sizeof(int), 1, f) == 1 &&
              fwrite(&c.t, sizeof(double), 1, f) == 1 && fwrite(&c.dt, sizeof(double), 1, f) == 1 &&
              fwrite(st.data(), sizeof(double), st.size(), f) == st.size() &&
              fwrite(wd.data(), sizeof(int), wd.size(), f) == wd.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? DGR_OK : DGR_ERR_ARG;
}

extern "C" dgr_status dgr_checkpoint_read(dgr_handle *h, int mi, const char *path)
{
    if (bad_member(h, mi) || !path) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[mi];
    FILE *f = fopen(path, "rb");
    if (!f) return DGR_ERR_ARG;
    char magic[8];
    int ne = 0;
    double t = 0, dt = 0;
    const size_t n3 = 3 * (size_t)M.ne;
    std::vector<double> st(3 * n3);
    std::vector<int> wd(M.ne);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "DGRCKPT1", 8) == 0 && fread(&ne, sizeof(int), 1, f) == 1 &&
              ne == M.ne && fread(&t, sizeof(double), 1, f) == 1 && fread(&dt, sizeof(double), 1, f) == 1 &&
              fread(st.data(), sizeof(double), st.size(), f) == st.size() &&
              fread(wd.data(), sizeof(int), wd.size(), f) == wd.size();
    fclose(f);
    if (!ok) return DGR_ERR_ARG;
    dgr_status r = dgr_member_set_state(h, mi, st.data(), st.data() + n3, st.data() + 2 * n3);
    if (r) return r;
    r = dgr_member_set_wd(h, mi, wd.data());
    if (r) return r;
    return dgr_member_set_time(h, mi, t, dt);
}

extern "C" dgr_status dgr_member_layout(dgr_handle *h, int mi, int *el_offset, int *num_el)
{
    if (bad_member(h, mi)) return DGR_ERR_ARG;
    if (el_offset) *el_offset = h->mem[mi].el_off;
    if (num_el) *num_el = h->mem[mi].ne;
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
    DGR_HIP(dalloc(&dout, 9 * (size_t)n));
    DGR_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
    DGR_HIP(hipMemcpy(dy, y, n * sizeof(double), hipMemcpyHostToDevice));
    launch_math_selftest(n, dx, dy, dout, 0);
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipMemcpy(out, dout, 9 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
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

// Times run steps (the path dgr_run takes) of each configuration that try()
// sets up, from the current state, which is restored after each (state, WD
// flags and step bookkeeping of every member); ms[i]: ms per step of
// configuration i.  All configurations compute the same bits.
template <typename Try>
static dgr_status time_configs(dgr_handle *h, int steps, int n, Try &&try_config, double *ms_out)
{
    const size_t P = h->pitch;
    double *w_save = nullptr;
    int *wd_save = nullptr;
    std::vector<Ctrl> c_save(h->nmem);
    DGR_HIP(dalloc(&w_save, 9 * P));
    DGR_HIP(dalloc(&wd_save, P));
    DGR_HIP(hipMemcpyAsync(w_save, h->d_w, 9 * P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(wd_save, h->d_wd, P * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
    DGR_HIP(hipMemcpyAsync(c_save.data(), h->d_ctrl, h->nmem * sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    hipEvent_t e0, e1;
    DGR_HIP(hipEventCreate(&e0));
    DGR_HIP(hipEventCreate(&e1));
    const int prof = h->profiling;
    h->profiling = 0;
    for (int i = 0; i < n; ++i) {
        try_config(i);
        // one untimed step to warm the code objects, then the timed steps
        enqueue_run_prologue(h);
        enqueue_run_step(h);
        DGR_HIP(hipEventRecord(e0, h->stream));
        for (int k = 0; k < steps; ++k) enqueue_run_step(h);
        DGR_HIP(hipEventRecord(e1, h->stream));
        DGR_HIP(hipEventSynchronize(e1));
        DGR_HIP(hipGetLastError());
        float ms = 0.f;
        DGR_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_out[i] = ms / steps;
        DGR_HIP(hipMemcpyAsync(h->d_w, w_save, 9 * P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        DGR_HIP(hipMemcpyAsync(h->d_wd, wd_save, P * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
        DGR_HIP(hipMemcpyAsync(h->d_ctrl, c_save.data(), h->nmem * sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
        DGR_HIP(hipStreamSynchronize(h->stream));
    }
    h->profiling = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(w_save);
    (void)hipFree(wd_save);
    return DGR_OK;
}

// Times `steps` run steps of every kernel variant from the current state and
// keeps the fastest.
extern "C" dgr_status dgr_autotune(dgr_handle *h, int steps, int *chosen, double *ms_per_step)
{
    if (!h || steps <= 0) return DGR_ERR_ARG;
    double ms[DGR_NUM_VARIANTS];
    const int old = h->variant;
    const dgr_status st = time_configs(h, steps, DGR_NUM_VARIANTS, [&](int v) { h->variant = v; }, ms);
    if (st != DGR_OK) {
        h->variant = old;
        return st;
    }
    int best = 0;
    for (int v = 0; v < DGR_NUM_VARIANTS; ++v) {
        if (ms_per_step) ms_per_step[v] = ms[v];
        if (ms[v] < ms[best]) best = v;
    }
    h->variant = best;
    if (chosen) *chosen = best;
    return DGR_OK;
}

extern "C" dgr_status dgr_set_block(dgr_handle *h, int threads)
{
    if (!h || (threads != 64 && threads != 128 && threads != 256)) return DGR_ERR_ARG;
    h->m.wg = threads;
    return DGR_OK;
}

extern "C" dgr_status dgr_get_block(dgr_handle *h, int *threads)
{
    if (!h || !threads) return DGR_ERR_ARG;
    *threads = h->m.wg;
    return DGR_OK;
}

extern "C" dgr_status dgr_autotune_block(dgr_handle *h, int steps, int *chosen, double *ms_per_step)
{
    if (!h || steps <= 0) return DGR_ERR_ARG;
    static const int sizes[3] = {64, 128, 256};
    double ms[3];
    const int old = h->m.wg;
    const dgr_status st = time_configs(h, steps, 3, [&](int i) { h->m.wg = sizes[i]; }, ms);
    if (st != DGR_OK) {
        h->m.wg = old;
        return st;
    }
    int best = 2;
    for (int i = 0; i < 3; ++i) {
        if (ms_per_step) ms_per_step[i] = ms[i];
        if (ms[i] < ms[best]) best = i;
    }
    h->m.wg = sizes[best];
    if (chosen) *chosen = sizes[best];
    return DGR_OK;
}
