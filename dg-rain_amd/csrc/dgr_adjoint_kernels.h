// dgr_adjoint_kernels.h -- the reverse sweep of the RK step on the GPU.
// Included by dgr_kernels.hip inside namespace dgr: the kernels re-evaluate
// the forward with its device functions (geometry_of, reflect, the member
// mapping) and differentiate with dgr_adjoint.h.
//
// One reverse RK stage, for every member at once (element kernels on the
// element workgroups, edge kernels on the edge workgroups):
//   k_node_arg      node min/max of the stage output's element means and the
//                   element that attains each (the limiter's active bounds)
//   k_adj_limit     PDop2D reverse (wet/dry runs, elements it changed), then
//                   limiter pass reverse: adjoint of the pre-limit output
//                   (own part) and of the node bounds at the element's vertices
//   k_adj_rhs       node sums of the bound adjoints, gathered by the element
//                   attaining each bound -> element means -> stage output; RK update
//                   and element RHS reverse: adjoint of the stage input, of
//                   Fhat, of dt (per-workgroup partial) and of nodal z, n
//   k_adj_edge      Roe flux reverse per edge, written to the two element
//                   slots the edge feeds (each slot has exactly one edge, so
//                   no atomics): adjoint of the stage input at the edge
//                   midpoints, of z_edge, of the tide value, of max |lambda|
//   k_adj_combine   adds the edge part to the stage input adjoint
// Then k_adj_scalar turns the dt and tide adjoints into the previous step's
// max-|lambda| adjoint (dt = cfl * minEdgLength / max|lambda|).
#pragma once

__device__ __forceinline__ void adj_scalar_member(const MeshDev &m, int mid, const AdjScalarArgs &sa, AdjCtrl *actl);

// limiter reverse.  U: pre-limit stage output; Sb (+ slotb when given): the
// adjoint of the limited (and PDop'd) output, slotb being the edge part left
// by the following stage's k_adj_edge, whose z_edge adjoints (zbe) are moved
// to the element's vertices here too.  With wet/dry on, PDop2D's reverse
// (pdop_adj, on the branch the forward took) turns that into the adjoint of
// the limited state first, and adds the adjoint of the vertex depths it
// reads to zvb.  Writes Ub (own part of the pre-limit adjoint) and
// vmb[18][pitch]: adjoint of Vmin (f*3+k) and Vmax (9+f*3+k) at vertex k.
__global__ __launch_bounds__(256, 3) void k_adj_limit(MeshDev m, PhysDev ph, const double *__restrict__ U,
                                                   const double *__restrict__ nodemm, const double *__restrict__ Sb,
                                                   const double *__restrict__ slotb, const double *__restrict__ zbe,
                                                   double *__restrict__ zvb, double *__restrict__ Ub,
                                                   double *__restrict__ vmb, AdjCtrl *actl, AdjScalarArgs sa)
{
    const ElemLane L = elem_lane(m);
    // the previous step's scalar pass, by each member's first element
    // workgroup (its inputs are complete: the launches before wrote them)
    if (sa.active && L.lb == m.mem[L.mid].eblk0) adj_scalar_member(m, L.mid, sa, actl);
    if (!L.valid) return;
    const int e = L.e, P = m.pitch, nn = m.nn;
    int nd[3];
    double zv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { nd[k] = m.ev[k * P + e]; zv[k] = m.zn[nd[k]]; }
    // the nodal z adjoints this element adds to, loaded with its inputs
    double zv0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) zv0[k] = zvb[k * P + e];
    double zadd[3] = {0, 0, 0};   // this element's additions to zvb
    if (zbe) {
        // z_edge = (z_a + z_b)/2 of local edge k
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double b = zbe[k * P + e];
            zadd[k] += 0.5 * b;
            zadd[(k + 1) % 3] += 0.5 * b;
        }
    }
    // limit_adj field by field: the zeta pass decides modify, then each
    // field's bounds are loaded, its pass replayed and reversed, and its
    // adjoints stored, so one field's data is live at a time.  A pass that
    // has already run keeps its decisions, packed in one word (dec), and is
    // not replayed: the zeta pass always, the momentum passes where PDop
    // acted
    double S[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) S[f] = U[f * P + e];
    // every field's bounds, loaded together with the zeta bounds (one round
    // of gathers instead of one per field)
    double Bmin[3][3], Bmax[3][3];
#pragma unroll
    for (int iv = 0; iv < 3; ++iv)
#pragma unroll
        for (int k = 0; k < 3; ++k) { Bmin[iv][k] = nodemm[iv * nn + nd[k]]; Bmax[iv][k] = nodemm[(3 + iv) * nn + nd[k]]; }
    double Vmin[3], Vmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { Vmin[k] = Bmin[0][k]; Vmax[k] = Bmax[0][k]; }
    // the zeta pass decides modify (its decisions are replayed again in the
    // field loop rather than kept live through the PDop reverse)
    int modify = 0;
    double lim[9];
    unsigned dec[3] = {0u, 0u, 0u};
    {
        LimFieldDec d;
        limit_field_forward(S, 0, Vmin, Vmax, d);
        dec[0] = lim_dec_pack(d);
        const double w1 = S[0], w2 = S[1], w3 = S[2];
        const double Ha = (w1 - w2 + w3) + zv[0], Hb_ = (w1 + w2 - w3) + zv[1], Hc = (-w1 + w2 + w3) + zv[2];
        if (Ha > ph.h0 && Hb_ > ph.h0 && Hc > ph.h0 && Vmin[0] + zv[0] > ph.h0 && Vmin[1] + zv[1] > ph.h0 &&
            Vmin[2] + zv[2] > ph.h0)
            modify = 1;
        if (d.Vfin[0] + zv[0] < ph.h0 || d.Vfin[1] + zv[1] < ph.h0 || d.Vfin[2] + zv[2] < ph.h0) modify = 0;
        lim[0] = S[0]; lim[1] = S[1]; lim[2] = S[2];
        if (modify) {
            lim[0] = (d.Vfin[0] + d.Vfin[1]) / 2;
            lim[1] = (d.Vfin[1] + d.Vfin[2]) / 2;
            lim[2] = (d.Vfin[2] + d.Vfin[0]) / 2;
        }
    }
    // where PDop acted, its reverse leaves the adjoint of the limited state in
    // Ub (read back field by field below, each slot before it is written)
    bool pdop = false;
    if (ph.wetdry) {
        // PDop2D (wetDry2D.c:66-256) changes an element iff a vertex depth it
        // forms from the limited state is at most H0
        const double bz0 = 0.5 * (zv[0] + zv[1]), bz1 = 0.5 * (zv[1] + zv[2]), bz2 = 0.5 * (zv[2] + zv[0]);
        const double mh0 = lim[0] + bz0, mh1 = lim[1] + bz1, mh2 = lim[2] + bz2;
        const double H0 = ph.h0;
        if (mh0 - mh1 + mh2 <= H0 || mh0 + mh1 - mh2 <= H0 || -mh0 + mh1 + mh2 <= H0) {
            pdop = true;
            // the limited momentum, then PDop's reverse on the limited state
#pragma unroll
            for (int iv = 1; iv < 3; ++iv) {
                if (modify) {
                    LimFieldDec dq;
                    limit_field_forward(S, iv, Bmin[iv], Bmax[iv], dq);
                    dec[iv] = lim_dec_pack(dq);
                    lim[3 * iv] = (dq.Vfin[0] + dq.Vfin[1]) / 2;
                    lim[3 * iv + 1] = (dq.Vfin[1] + dq.Vfin[2]) / 2;
                    lim[3 * iv + 2] = (dq.Vfin[2] + dq.Vfin[0]) / 2;
                } else {
#pragma unroll
                    for (int k = 0; k < 3; ++k) lim[3 * iv + k] = S[3 * iv + k];
                }
            }
            double sb[9], lb[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) sb[f] = Sb[f * P + e] + (slotb ? slotb[f * P + e] : 0.0);
            pdop_adj(lim, zv, H0, ph.velzero, sb, lb, zadd);
#pragma unroll
            for (int f = 0; f < 9; ++f) Ub[f * P + e] = lb[f];
        }
    }
    if (zbe || pdop) {
#pragma unroll
        for (int k = 0; k < 3; ++k) zvb[k * P + e] = zv0[k] + zadd[k];
    }
    auto sb_of = [&](int f) { return pdop ? Ub[f * P + e] : Sb[f * P + e] + (slotb ? slotb[f * P + e] : 0.0); };
#pragma unroll
    for (int iv = 0; iv < 3; ++iv) {
        double sb[9], ub[9], Vminb[3] = {0, 0, 0}, Vmaxb[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) { sb[3 * iv + k] = sb_of(3 * iv + k); ub[3 * iv + k] = 0.0; }
        if (!modify) {
#pragma unroll
            for (int k = 0; k < 3; ++k) ub[3 * iv + k] += sb[3 * iv + k];
        } else {
            LimFieldDec d;
            if (iv == 0 || pdop) {
                d = lim_dec_unpack(dec[iv]);
            } else {
                limit_field_forward(S, iv, Bmin[iv], Bmax[iv], d);
            }
            limit_field_reverse(d, iv, sb, ub, Vminb, Vmaxb);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Ub[(3 * iv + k) * P + e] = ub[3 * iv + k];
            vmb[(iv * 3 + k) * P + e] = Vminb[k];
            vmb[(9 + iv * 3 + k) * P + e] = Vmaxb[k];
        }
    }
}

// node sums of the bound adjoints: for every node, the sums of vmb (the
// limiter reverse's bound adjoints, [18][pitch]) over the node's elements in
// n2e order, as nsum[8 i + f], f = 0..2 the min bounds of the three fields,
// 3..5 the max bounds (6..7 unused).  Only the element that attains a bound
// reads its sum (k_adj_rhs).  Entries are taken four at a time with clamped
// indices, so every load of a group is issued before the first add.
__global__ __launch_bounds__(256) void k_adj_node_sums(MeshDev m, const double *__restrict__ vmb,
                                                       double *__restrict__ nsum)
{
    const int i = xcd_block() * 256 + (int)threadIdx.x;
    if (i >= m.nn) return;
    const int P = m.pitch;
    double s[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) s[f] = 0.0;
    const int b = m.n2e_off[i], en = m.n2e_off[i + 1];
    for (int j = b; j < en; j += 4) {
        int v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = m.n2e[min(j + q, en - 1)];
        double x[4][6];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int el = v[q] & N2E_EL_MASK, k = (unsigned)v[q] >> N2E_K_SHIFT;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                x[q][f] = vmb[(f * 3 + k) * P + el];
                x[q][3 + f] = vmb[(9 + f * 3 + k) * P + el];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool in = j + q < en;
#pragma unroll
            for (int f = 0; f < 6; ++f) s[f] = in ? s[f] + x[q][f] : s[f];
        }
    }
    double2 *o = reinterpret_cast<double2 *>(nsum) + 4 * (size_t)i;
    o[0] = make_double2(s[0], s[1]);
    o[1] = make_double2(s[2], s[3]);
    o[2] = make_double2(s[4], s[5]);
}

// sum over a workgroup of 256 threads, valid in thread 0
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// RK update + element RHS reverse.
//   MODE 1: U = W + R dt             -> Wb += Ub + RHS'(W)^T (dt Ub)
//   MODE 2: U = (W + B) 0.5 + R dt/2 -> Inb = 0.5 Ub + RHS'(W)^T (dt/2 Ub), Bb = 0.5 Ub
// W: stage input, B: step-start state (MODE 2), U: pre-limit output, Ubown:
// own part of the pre-limit adjoint; the node-bound part is added here, from
// the node sums nsum of the limiter reverse's bound adjoints (k_adj_node_sums),
// for the elements that attain a bound.  Outputs: Inb (MODE 2: the
// stage-input adjoint, written; MODE 1: added to Bb, which is the step-start
// adjoint), Fhb (adjoint of Fhat), zvb/nvb (nodal z/n adjoints per element
// vertex, added), dtparts (this workgroup's dt adjoint).
template <int MODE>
__global__ __launch_bounds__(256, 2) void k_adj_rhs(MeshDev m, PhysDev ph, const double *__restrict__ W,
                                                    const double *__restrict__ B, const double *__restrict__ U,
                                                    const double *__restrict__ Ubown, const double *__restrict__ nsum,
                                                    const int *__restrict__ nodearg, const int *__restrict__ wd,
                                                    const StepRec *__restrict__ rec, double *Inb, double *Bb,
                                                    double *__restrict__ Fhb, double *__restrict__ zvb,
                                                    double *__restrict__ nvb, double *__restrict__ dtparts)
{
    const ElemLane L = elem_lane(m);
    double dtb = 0.0;
    if (L.valid) {
        const int e = L.e, P = m.pitch, nn = m.nn;
        const double dt = rec[L.mid].dt;
        double w[9], u[9], ub[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) { w[f] = W[f * P + e]; u[f] = U[f * P + e]; ub[f] = Ubown[f * P + e]; }
        // the element mean enters the node bounds it attains; a bound's
        // adjoint is the sum of the bound adjoints over the node's elements
        // (nsum), read here by the element that attains it
        int nd[3];
        bool own[3][6];
#pragma unroll
        for (int k = 0; k < 3; ++k) nd[k] = m.ev[k * P + e];
        // the vertex records (into this thread's LDS slots) and the nodal
        // adjoints this element adds to, in the same round of loads as the
        // attaining-element records
        __shared__ double4 vtx[3][256];
#pragma unroll
        for (int k = 0; k < 3; ++k) vtx[k][threadIdx.x] = m.nodes[nd[k]];
        double zv0[3], nv0[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { zv0[k] = zvb[k * P + e]; nv0[k] = nvb[k * P + e]; }
        // MODE 1 adds to the step-start adjoint: its current value waits in
        // this thread's LDS slots instead of being loaded at the end
        __shared__ double bbl[MODE == 1 ? 9 : 1][256];
        if (MODE == 1) {
#pragma unroll
            for (int f = 0; f < 9; ++f) bbl[f][threadIdx.x] = Bb[f * P + e];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int4 *na = reinterpret_cast<const int4 *>(nodearg) + 2 * nd[k];
            const int4 a0 = na[0], a1 = na[1];
            own[k][0] = a0.x == e; own[k][1] = a0.y == e; own[k][2] = a0.z == e;
            own[k][3] = a0.w == e; own[k][4] = a1.x == e; own[k][5] = a1.y == e;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bool any = false;
#pragma unroll
            for (int f = 0; f < 6; ++f) any |= own[k][f];
            if (!any) continue;
            const double2 *ns = reinterpret_cast<const double2 *>(nsum) + 4 * (size_t)nd[k];
            const double2 s01 = ns[0], s23 = ns[1], s45 = ns[2];
            const double sb[6] = {s01.x, s01.y, s23.x, s23.y, s45.x, s45.y};
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                double add = 0.0;
                if (own[k][f]) add += sb[f];
                if (own[k][3 + f]) add += sb[3 + f];
                add *= 1.0 / 3.0;
                ub[3 * f] += add;
                ub[3 * f + 1] += add;
                ub[3 * f + 2] += add;
            }
        }
        // the RHS the forward used, recovered from the update (it only
        // weighs the dt adjoint, so reciprocals do)
        double Rb[9], R[9];
        const double rdt = 1.0 / dt;
        if (MODE == 1) {
#pragma unroll
            for (int f = 0; f < 9; ++f) { R[f] = (u[f] - w[f]) * rdt; Rb[f] = dt * ub[f]; dtb += R[f] * ub[f]; }
        } else {
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                const double bf = B[f * P + e];
                R[f] = (u[f] - (w[f] + bf) * 0.5) * (2.0 * rdt);
                Rb[f] = 0.5 * dt * ub[f];
                dtb += 0.5 * R[f] * ub[f];
            }
        }
        // geometry enters derivatives only: reciprocal products.  Evaluated
        // once for the surface part and again for the volume part, so only
        // one part's geometry is live at a time; the vertex data both
        // evaluations and the nodal adjoints need come from this thread's
        // own LDS slots
        {
            double4 Pv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) Pv[k] = vtx[k][threadIdx.x];
            GeoRegs G;
            DivRecip D;
            geometry_of_vertices(m, e, Pv, ph.g, [&](int i, double v) { G.v[i] = v; }, D);
            double Fb[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) Fb[f] = 0.0;
            element_rhs_adj_surface(G, Rb, Fb);
#pragma unroll
            for (int f = 0; f < 9; ++f) Fhb[f * P + e] = Fb[f];
        }
        // (the compiler may not carry the first evaluation's vertex values
        // across: they come back from LDS)
        asm volatile("" ::: "memory");
        double4 Pw[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) Pw[k] = vtx[k][threadIdx.x];
        GeoRegs G;
        DivRecip D;
        geometry_of_vertices(m, e, Pw, ph.g, [&](int i, double v) { G.v[i] = v; }, D);
        double Wb[9], mzb[3] = {0, 0, 0}, gnnb[3] = {0, 0, 0}, bgxb = 0.0, bgyb = 0.0;
#pragma unroll
        for (int f = 0; f < 9; ++f) Wb[f] = 0.0;
        element_rhs_adj_volume(G, w, wd[e], ph.g, Rb, Wb, mzb, gnnb, bgxb, bgyb);
        if (MODE == 1) {
#pragma unroll
            for (int f = 0; f < 9; ++f) Bb[f * P + e] = bbl[f][threadIdx.x] + (ub[f] + Wb[f]);
        } else {
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                Inb[f * P + e] = 0.5 * ub[f] + Wb[f];
                Bb[f * P + e] = 0.5 * ub[f];   // first write of the step-start adjoint
            }
        }
        // geometry -> nodal z and n.  bgx = g |jac| weight sum_k mz_k gP_k,x
        const double cbed = ph.g * G(GX_AJ) * (1.0 / 6);
#pragma unroll
        for (int k = 0; k < 3; ++k) mzb[k] += cbed * (G(GX_GP + 2 * k) * bgxb + G(GX_GP + 2 * k + 1) * bgyb);
        asm volatile("" ::: "memory");
        double4 Pv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) Pv[k] = vtx[k][threadIdx.x];
        double zb_[3] = {0, 0, 0}, nb_[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int k1 = (k + 1) % 3;
            // mz_k = 0.5 (z_k + z_k1), gnn_k = g mn_k^2, mn_k = 0.5 (n_k + n_k1)
            zb_[k] += 0.5 * mzb[k];
            zb_[k1] += 0.5 * mzb[k];
            const double mn = 0.5 * (Pv[k].w + Pv[k1].w);
            nb_[k] += ph.g * mn * gnnb[k];
            nb_[k1] += ph.g * mn * gnnb[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) { zvb[k * P + e] = zv0[k] + zb_[k]; nvb[k * P + e] = nv0[k] + nb_[k]; }
    }
    const double s = block_sum(dtb);
    if (threadIdx.x == 0) dtparts[L.lb] = s;
}

// wall reflection (compute2DL.c:179-189) as a 2x2 map r = A q
__device__ __forceinline__ void reflect_matrix(double nx, double ny, double dn, double A[2][2])
{
    const double tx = -ny, ty = nx;
    // rx = (ty*(-qn) - ny*qt) dn, ry = (-tx*(-qn) + nx*qt) dn, qn = q.n, qt = q.t
    A[0][0] = (-ty * nx - ny * tx) * dn;
    A[0][1] = (-ty * ny - ny * ty) * dn;
    A[1][0] = (tx * nx + nx * tx) * dn;
    A[1][1] = (tx * ny + nx * ty) * dn;
}

// the edge's inputs, as the forward builds them (compute2DL.c:98-199)
struct EdgeState {
    int el_in, el_ex, kin, kex, inf, code;
    bool boundary, prescribed, tide, zex_from_in, qn_from_in;
    double nx, ny, dn, zb;
    double zin, qxin, qyin, zex, qxex, qyex;
};

__device__ __forceinline__ EdgeState edge_state(const MeshDev &m, const MemberDev &M, const StepRec &rc,
                                                const double *__restrict__ win, const double *__restrict__ bzeta,
                                                const double *__restrict__ bqn, int ed)
{
    EdgeState s;
    const int P = m.pitch;
    const int4 E = m.edges[ed];
    s.inf = m.einfo[ed];
    s.el_in = E.x;
    s.el_ex = E.y;
    s.kin = s.inf & ED_KIN_MASK;
    s.kex = (s.inf >> ED_KEX_SHIFT) & 3;
    s.boundary = (s.inf & EI_BOUNDARY) != 0;
    s.prescribed = (s.inf & EI_PRESCRIBED) != 0;
    s.code = (s.inf >> EI_CODE_SHIFT) & 0xff;
    const double4 G = m.egeo[ed];   // the forward's edge geometry (k_edge_geometry)
    s.nx = G.x;
    s.ny = G.y;
    const double tx = -s.ny, ty = s.nx;
    s.zb = G.z;
    s.dn = G.w;
    s.zin = win[s.kin * P + s.el_in];
    s.qxin = win[(3 + s.kin) * P + s.el_in];
    s.qyin = win[(6 + s.kin) * P + s.el_in];
    s.tide = false;
    s.zex_from_in = false;
    s.qn_from_in = false;
    if (s.boundary) { s.zex = s.zin; s.qxex = s.qxin; s.qyex = s.qyin; }
    else {
        s.zex = win[s.kex * P + s.el_ex];
        s.qxex = win[(3 + s.kex) * P + s.el_ex];
        s.qyex = win[(6 + s.kex) * P + s.el_ex];
    }
    if (s.prescribed) {
        s.tide = M.tide_code != 0 && s.code == M.tide_code;
        s.zex = s.tide ? rc.tide : bzeta[ed];
        double qnex = s.tide ? 1000000.0 : bqn[ed];
        if (M.sentinels && s.zex < 0) { s.zex = s.zin; s.zex_from_in = true; }
        if (M.sentinels && qnex > 1000) { qnex = s.qxin * s.nx + s.qyin * s.ny; s.qn_from_in = true; }
        const double qtex = s.qxin * tx + s.qyin * ty;
        s.qxex = (ty * qnex - s.ny * qtex) * s.dn;
        s.qyex = (-tx * qnex + s.nx * qtex) * s.dn;
    } else if (s.boundary) {
        reflect(s.qxin, s.qyin, s.nx, s.ny, s.dn, s.qxex, s.qyex);
        s.zex_from_in = true;
    }
    return s;
}

// the |lambda| operands of an edge (its own positions 2*edge, 2*edge+1)
__device__ __forceinline__ int edge_lams(const EdgeState &s, const int *__restrict__ wd, bool wetdry, double g,
                                         double lam[2])
{
    double F[3];
    RoeIn a;
    if (wetdry) {
        const int wd_in = wd[s.el_in], wd_ex = s.boundary ? wd_in : wd[s.el_ex];
        if (wd_in == 0 && wd_ex == 0) {
            double rx, ry, sx, sy;
            reflect(s.qxin, s.qyin, s.nx, s.ny, s.dn, rx, ry);
            a = {s.zin, s.zin, s.qxin, rx, s.qyin, ry, s.zb};
            lam[0] = roe_eval(F, a, s.nx, s.ny, 0.0, g);
            reflect(s.qxex, s.qyex, s.nx, s.ny, s.dn, sx, sy);
            a = {s.zex, s.zex, sx, s.qxex, sy, s.qyex, s.zb};
            lam[1] = roe_eval(F, a, s.nx, s.ny, 0.0, g);
            return 2;
        }
    }
    a = {s.zin, s.zex, s.qxin, s.qxex, s.qyin, s.qyex, s.zb};
    lam[0] = roe_eval(F, a, s.nx, s.ny, g, g);
    return 1;
}

// count per member the edges whose |lambda| operand equals the recorded max
__global__ __launch_bounds__(256) void k_adj_ties(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                                  const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                  const int *__restrict__ wd, const StepRec *__restrict__ rec_lam,
                                                  const StepRec *__restrict__ rec, AdjCtrl *actl)
{
    const int lb = xcd_block();
    const int mid = m.dblk_mem[lb];
    const MemberDev &M = m.mem[mid];
    const int ed = lb * 256 + (int)threadIdx.x;
    if (ed >= M.ed_off + M.ned) return;
    if (actl[mid].lamb == 0.0) return;
    const EdgeState s = edge_state(m, M, rec[mid], win, bzeta, bqn, ed);
    double lam[2];
    const int n = edge_lams(s, wd, ph.wetdry != 0, ph.g, lam);
    const double target = __longlong_as_double((long long)rec_lam[mid].lam2);
    int c = 0;
    for (int i = 0; i < n; ++i) c += lam[i] == target;
    if (c) atomicAdd(&actl[mid].ntie, c);
}

// Roe flux reverse of one edge.  W: stage input; Fhb: adjoint of the Fhat
// slots.  Writes slotb[9][pitch] (adjoint of the stage input at the two
// slots the edge feeds; a boundary edge writes its one slot) and zbe[3][pitch]
// (the z_edge adjoint in el_in's slot, 0 in el_ex's) and returns the edge's
// tide-value adjoint.  With lamb != 0 the edges whose |lambda| operand equals
// the recorded max (rec_lam) share lamb.  WETDRY false: the edge's sides are
// wet (or wet/dry is off): one Roe evaluation.  GATE: the edge's wet/dry
// flags are loaded with its other inputs, and an edge with a side not wet
// is left alone (*dry set, nothing written).
template <bool WETDRY, bool GATE = false>
__device__ __forceinline__ double adj_edge_one(const MeshDev &m, const PhysDev &ph, const MemberDev &M, int mid,
                                               int ed, const double *__restrict__ win,
                                               const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                               const int *__restrict__ wd, const double *__restrict__ Fhb,
                                               const StepRec *__restrict__ rec, const StepRec *__restrict__ rec_lam,
                                               const AdjCtrl *__restrict__ actl, int use_lambda,
                                               double *__restrict__ slotb, double *__restrict__ zbe,
                                               bool *dry = nullptr)
{
    double tideb = 0.0;
    const int P = m.pitch;
    const double g = ph.g;
    const EdgeState s = edge_state(m, M, rec[mid], win, bzeta, bqn, ed);
    double lam_share = 0.0;
    const AdjCtrl &ac = actl[mid];
    // the tie count: recorded by the taped forward (use_lambda 2) or
    // counted by k_adj_ties (1)
    const int ntie = (use_lambda == 2) ? rec_lam[mid].ntie : ac.ntie;
    if (use_lambda && ac.lamb != 0.0 && ntie > 0) lam_share = ac.lamb / ntie;
    const double target = __longlong_as_double((long long)rec_lam[mid].lam2);
    double Fin_b[3], Fex_b[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        Fin_b[f] = Fhb[(f * 3 + s.kin) * P + s.el_in];
        Fex_b[f] = s.boundary ? 0.0 : -Fhb[(f * 3 + s.kex) * P + s.el_ex];
    }
    if (GATE) {
        // both sides wet (a boundary edge's outside is its inside)
        const int wi = wd[s.el_in];
        if (!(wi == 1 && (s.boundary || wd[s.el_ex] == 1))) {
            *dry = true;
            return 0.0;
        }
    }
    // adjoints of the edge-level inputs.  Up to two Roe evaluations per
    // edge: (wet) one with g; (one side dry) the g flux and the lg = 0
    // flux of the same states; (both dry) el_in's and el_ex's reflection
    // fluxes.  One roe_eval/roe_adj in a loop keeps a single copy of the
    // reverse code (and its registers) for every case.
    double zinb = 0, qxinb = 0, qyinb = 0, zexb = 0, qxexb = 0, qyexb = 0, zbb = 0;
    double Ar[2][2];
    reflect_matrix(s.nx, s.ny, s.dn, Ar);
    int mode = 0;   // 0 wet, 1 one side dry, 2 both dry
    bool in_wet = true, ex_wet = true;
    if (WETDRY) {
        const int wd_in = wd[s.el_in], wd_ex = s.boundary ? wd_in : wd[s.el_ex];
        in_wet = wd_in == 1;
        ex_wet = s.boundary ? in_wet : (wd_ex == 1);
        if (wd_in == 0 && wd_ex == 0) mode = 2;
        else if (!(in_wet && ex_wet)) mode = 1;
    }
    RoeIn acc = {0, 0, 0, 0, 0, 0, 0};   // identity-mapped input adjoints (modes 0, 1)
    const int nev = (mode == 0) ? 1 : 2;
#pragma unroll 1
    for (int j = 0; j < nev; ++j) {
        RoeIn a;
        double lg = g, Fb[3];
        bool lam_op = true;
        if (mode == 2) {
            lg = 0.0;
            if (j == 0) {
                double rx, ry;
                reflect(s.qxin, s.qyin, s.nx, s.ny, s.dn, rx, ry);
                a = RoeIn{s.zin, s.zin, s.qxin, rx, s.qyin, ry, s.zb};
            } else {
                double sx, sy;
                reflect(s.qxex, s.qyex, s.nx, s.ny, s.dn, sx, sy);
                a = RoeIn{s.zex, s.zex, sx, s.qxex, sy, s.qyex, s.zb};
            }
#pragma unroll
            for (int f = 0; f < 3; ++f) Fb[f] = (j == 0) ? Fin_b[f] : Fex_b[f];
        } else {
            a = RoeIn{s.zin, s.zex, s.qxin, s.qxex, s.qyin, s.qyex, s.zb};
            if (mode == 0) {
#pragma unroll
                for (int f = 0; f < 3; ++f) Fb[f] = Fin_b[f] + Fex_b[f];
            } else if (j == 0) {   // the g flux feeds the wet sides
#pragma unroll
                for (int f = 0; f < 3; ++f) Fb[f] = (in_wet ? Fin_b[f] : 0.0) + (ex_wet ? Fex_b[f] : 0.0);
            } else {               // the lg = 0 flux the dry sides (|lambda| counted once)
                lg = 0.0;
                lam_op = false;
#pragma unroll
                for (int f = 0; f < 3; ++f) Fb[f] = (in_wet ? 0.0 : Fin_b[f]) + (ex_wet ? 0.0 : Fex_b[f]);
            }
        }
        // (|lambda| does not depend on lg: the tie test inside roe_adj_tie
        // sees the operands the forward folded)
        RoeIn ib = (mode == 2) ? RoeIn{0, 0, 0, 0, 0, 0, 0} : acc;
        roe_adj_tie(a, s.nx, s.ny, lg, g, Fb, lam_op ? lam_share : 0.0, target, ib);
        if (mode != 2) {
            acc = ib;
        } else if (j == 0) {
            zinb += ib.zi + ib.ze;
            qxinb += ib.qxi + Ar[0][0] * ib.qxe + Ar[1][0] * ib.qye;
            qyinb += ib.qyi + Ar[0][1] * ib.qxe + Ar[1][1] * ib.qye;
            zbb += ib.zb;
        } else {
            zexb += ib.zi + ib.ze;
            qxexb += ib.qxe + Ar[0][0] * ib.qxi + Ar[1][0] * ib.qyi;
            qyexb += ib.qye + Ar[0][1] * ib.qxi + Ar[1][1] * ib.qyi;
            zbb += ib.zb;
        }
    }
    if (mode != 2) {
        zinb += acc.zi; zexb += acc.ze; qxinb += acc.qxi; qxexb += acc.qxe;
        qyinb += acc.qyi; qyexb += acc.qye; zbb += acc.zb;
    }
    // boundary data -> slots (compute2DL.c:149-199)
    double exz = 0, exqx = 0, exqy = 0;   // adjoint of the el_ex slot
    if (s.prescribed) {
        if (s.zex_from_in) zinb += zexb;
        else if (s.tide) tideb += zexb;
        const double tx = -s.ny, ty = s.nx;
        const double ax = s.qn_from_in ? s.nx : 0.0, ay = s.qn_from_in ? s.ny : 0.0;
        // qxex = (ty qn - ny qt) dn, qyex = (-tx qn + nx qt) dn, qt = qxin tx + qyin ty
        qxinb += ((ty * ax - s.ny * tx) * qxexb + (-tx * ax + s.nx * tx) * qyexb) * s.dn;
        qyinb += ((ty * ay - s.ny * ty) * qxexb + (-tx * ay + s.nx * ty) * qyexb) * s.dn;
    } else if (s.boundary) {
        zinb += zexb;
        qxinb += Ar[0][0] * qxexb + Ar[1][0] * qyexb;
        qyinb += Ar[0][1] * qxexb + Ar[1][1] * qyexb;
    } else {
        exz = zexb; exqx = qxexb; exqy = qyexb;
    }
    slotb[s.kin * P + s.el_in] = zinb;
    slotb[(3 + s.kin) * P + s.el_in] = qxinb;
    slotb[(6 + s.kin) * P + s.el_in] = qyinb;
    zbe[s.kin * P + s.el_in] = zbb;
    if (!s.boundary) {
        slotb[s.kex * P + s.el_ex] = exz;
        slotb[(3 + s.kex) * P + s.el_ex] = exqx;
        slotb[(6 + s.kex) * P + s.el_ex] = exqy;
        zbe[s.kex * P + s.el_ex] = 0.0;
    }
    return tideb;
}

// every edge in one launch (WETDRY false: wet/dry off, one Roe evaluation per
// edge, 3 waves/SIMD; true: up to two, 2 waves/SIMD)
template <bool WETDRY>
__global__ __launch_bounds__(256, WETDRY ? 2 : 3) void k_adj_edge(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                                  const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                  const int *__restrict__ wd, const double *__restrict__ Fhb,
                                                  const StepRec *__restrict__ rec, const StepRec *__restrict__ rec_lam,
                                                  const AdjCtrl *__restrict__ actl, int use_lambda,
                                                  double *__restrict__ slotb, double *__restrict__ zbe,
                                                  double *__restrict__ tideparts)
{
    const int lb = xcd_block();
    const int mid = m.dblk_mem[lb];
    const MemberDev &M = m.mem[mid];
    const int ed = lb * 256 + (int)threadIdx.x;
    double tideb = 0.0;
    if (ed < M.ed_off + M.ned)
        tideb = adj_edge_one<WETDRY>(m, ph, M, mid, ed, win, bzeta, bqn, wd, Fhb, rec, rec_lam, actl, use_lambda,
                                     slotb, zbe);
    const double t = block_sum(tideb);
    if (threadIdx.x == 0) tideparts[lb] = t;
}

// On large meshes with wet/dry the reverse runs in two launches instead.
// k_adj_edge_wet reverses the edges whose sides are both wet (one Roe
// evaluation, 3 waves/SIMD), lists each workgroup's other edges in thread
// order (blist[256 lb + i]) and queues the workgroups that have any ({lb,
// count}, one atomic per such workgroup), keeping their wet edges' tide
// adjoints (tb[ed]).  k_adj_edge_dry (up to two evaluations per edge, the
// register budget of 2 waves/SIMD) takes the queued workgroups on a small
// grid, one workgroup each, and forms their tide partial over all 256 edges
// in thread order, as k_adj_edge<true> does: same bits, whatever the queue
// order.
__global__ __launch_bounds__(256, 3) void k_adj_edge_wet(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                                      const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                      const int *__restrict__ wd, const double *__restrict__ Fhb,
                                                      const StepRec *__restrict__ rec,
                                                      const StepRec *__restrict__ rec_lam,
                                                      const AdjCtrl *__restrict__ actl, int use_lambda,
                                                      double *__restrict__ slotb, double *__restrict__ zbe,
                                                      double *__restrict__ tideparts, int *__restrict__ blist,
                                                      double *__restrict__ tb, int2 *__restrict__ queue,
                                                      int *qcount)
{
    __shared__ int wcnt[4];
    const int lb = xcd_block();
    const int mid = m.dblk_mem[lb];
    const MemberDev &M = m.mem[mid];
    const int ed = lb * 256 + (int)threadIdx.x;
    double tideb = 0.0;
    bool dry = false;
    if (ed < M.ed_off + M.ned)
        tideb = adj_edge_one<false, true>(m, ph, M, mid, ed, win, bzeta, bqn, wd, Fhb, rec, rec_lam, actl,
                                          use_lambda, slotb, zbe, &dry);
    const unsigned long long bal = __ballot(dry);
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) base += (i < wv) ? wcnt[i] : 0;
    const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    if (dry) blist[lb * 256 + base + __popcll(bal & ((1ull << lane) - 1))] = ed;
    if (tot) tb[ed] = tideb;
    const double t = block_sum(tideb);
    if (threadIdx.x == 0) {
        tideparts[lb] = t;
        if (tot) queue[atomicAdd(qcount, 1)] = make_int2(lb, tot);
    }
}

__global__ __launch_bounds__(256, 2) void k_adj_edge_dry(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                                      const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                      const int *__restrict__ wd, const double *__restrict__ Fhb,
                                                      const StepRec *__restrict__ rec,
                                                      const StepRec *__restrict__ rec_lam,
                                                      const AdjCtrl *__restrict__ actl, int use_lambda,
                                                      double *__restrict__ slotb, double *__restrict__ zbe,
                                                      double *__restrict__ tideparts, const int *__restrict__ blist,
                                                      const double *__restrict__ tb, const int2 *__restrict__ queue,
                                                      const int *qcount, int *qnext)
{
    __shared__ double tbl[256];
    // the counter of the next k_adj_edge_wet, unused by this launch
    if (blockIdx.x == 0 && threadIdx.x == 0) *qnext = 0;
    const int nq = *qcount;
    for (int q = (int)blockIdx.x; q < nq; q += (int)gridDim.x) {
        const int2 Q = queue[q];
        const int b = Q.x;
        const int mid = m.dblk_mem[b];
        const MemberDev &M = m.mem[mid];
        tbl[threadIdx.x] = tb[b * 256 + (int)threadIdx.x];
        __syncthreads();
        if ((int)threadIdx.x < Q.y) {
            const int ed = blist[b * 256 + (int)threadIdx.x];
            tbl[ed - b * 256] = adj_edge_one<true>(m, ph, M, mid, ed, win, bzeta, bqn, wd, Fhb, rec, rec_lam, actl,
                                                   use_lambda, slotb, zbe);
        }
        __syncthreads();
        const double t = block_sum(tbl[threadIdx.x]);
        if (threadIdx.x == 0) tideparts[b] = t;
        __syncthreads();   // tbl and block_sum's shared slots are reused next round
    }
}

// Dst += slotb (stage input adjoint); the z_edge adjoints to the edge's two
// vertices (z_edge = (z_a + z_b)/2)
__global__ __launch_bounds__(256) void k_adj_combine(MeshDev m, double *Dst, const double *__restrict__ slotb,
                                                     const double *__restrict__ zbe, double *__restrict__ zvb)
{
    const ElemLane L = elem_lane(m);
    if (!L.valid) return;
    const int e = L.e, P = m.pitch;
#pragma unroll
    for (int f = 0; f < 9; ++f) Dst[f * P + e] += slotb[f * P + e];
    double z[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double b = zbe[k * P + e];
        z[k] += 0.5 * b;
        z[(k + 1) % 3] += 0.5 * b;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) zvb[k * P + e] += z[k];
}

// per member, after the two stages of step n are reversed: the dt adjoint of
// step n (workgroup partials of both stages plus the start-time term), the
// tide time derivative, and the adjoint of the previous step's max |lambda|
// (dt_n = cfl minEdgLength / lambda_{n-1} unless the final-time clamp acted).
__device__ __forceinline__ void adj_scalar_member(const MeshDev &m, int mid, const AdjScalarArgs &sa, AdjCtrl *actl)
{
    const MemberDev &M = m.mem[mid];
    const double *dtparts = sa.dtparts, *tideparts = sa.tideparts;
    const int maxparts = sa.maxparts;
    __shared__ double red[2][4];
    double dtb = 0.0, tb = 0.0;
    for (int i = threadIdx.x; i < M.eblks; i += 256)
        dtb += dtparts[M.eblk0 + i] + dtparts[maxparts + M.eblk0 + i];
    for (int i = threadIdx.x; i < M.dblks; i += 256)
        tb += tideparts[M.dblk0 + i] + tideparts[maxparts + M.dblk0 + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { dtb += __shfl_xor(dtb, o, 64); tb += __shfl_xor(tb, o, 64); }
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = dtb; red[1][threadIdx.x >> 6] = tb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        dtb = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        tb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        AdjCtrl &a = actl[mid];
        const StepRec r = sa.rec[mid];
        dtb += a.S;   // t_k = t_0 + sum_{j<k} dt_j
        double tbar = 0.0;
        if (M.tide_code) {
            const double w = 6.283185307179586 / M.tide_period;
            tbar = tb * M.tide_amp * w * cos(w * r.t + M.tide_phase);
        }
        a.S += tbar;
        // the max-|lambda| adjoint this step's edges shared needs a recorded tie
        // count; a fold that saw a NaN operand has none
        if (sa.rec_ties && a.lamb != 0.0 && r.ntie <= 0) a.unsupported = 1;
        a.lamb = 0.0;
        a.ntie = 0;
        if (sa.has_prev) {
            const double lam = __longlong_as_double((long long)sa.rec_prev[mid].lam2);
            if (r.dt == M.cfl * M.min_edge / lam) a.lamb = -r.dt / lam * dtb;
        }
    }
    __syncthreads();
}

__global__ void k_adj_scalar(MeshDev m, AdjScalarArgs sa, AdjCtrl *actl)
{
    adj_scalar_member(m, blockIdx.x, sa, actl);
}

// loss gradient of the gauges sampled at this state: h_g = mean over the
// element's midpoints of zeta + z_mid, loss = sum (h_g - obs)^2; one thread
// per (member, gauge), atomic adds because a gauge element may repeat
__global__ void k_adj_gauges(MeshDev m, int ngauge, const int *__restrict__ gauge_el,
                             const double *__restrict__ hval, const double *__restrict__ obs, double *Wb, double *zvb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nmem * ngauge) return;
    const int mid = t / ngauge;
    const MemberDev &M = m.mem[mid];
    const int P = m.pitch;
    const int e = M.el_off + gauge_el[t];
    const double w = 2.0 * (hval[t] - obs[t]) / 3.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicAdd(&Wb[k * P + e], w);
        atomicAdd(&zvb[k * P + e], w);   // z_mid_k = (z_k + z_k1)/2: each vertex in two midpoints
    }
}

// forward sample: h at the gauges, one thread per (member, gauge), member
// mid's gauge g at hval[mid * mstride + g]; with obs the squared misfits are
// summed per member in gauge order
__global__ void k_gauge_sample(MeshDev m, int ngauge, const int *__restrict__ gauge_el, const double *__restrict__ w,
                               double *__restrict__ hval, int mstride)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nmem * ngauge) return;
    const MemberDev &M = m.mem[t / ngauge];
    const int P = m.pitch;
    const int e = M.el_off + gauge_el[t];
    double h = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double za = m.zn[m.ev[k * P + e]], zb = m.zn[m.ev[((k + 1) % 3) * P + e]];
        h += w[k * P + e] + 0.5 * (za + zb);
    }
    hval[(t / ngauge) * mstride + t % ngauge] = h / 3.0;
}

__global__ void k_gauge_loss(MeshDev m, int ngauge, const double *__restrict__ hval, const double *__restrict__ obs,
                             AdjCtrl *actl)
{
    const int mid = blockIdx.x * 64 + threadIdx.x;
    if (mid >= m.nmem) return;
    double loss = 0.0;
    for (int g = 0; g < ngauge; ++g) {
        const double r = hval[mid * ngauge + g] - obs[mid * ngauge + g];
        loss += r * r;
    }
    actl[mid].loss += loss;
}

// step record: at the start (t, dt, tide) or the end (max |lambda| of stage 2)
__global__ void k_step_record(MeshDev m, const Ctrl *ctrls, StepRec *rec, int at_end)
{
    const int mid = blockIdx.x * 64 + threadIdx.x;
    if (mid >= m.nmem) return;
    const Ctrl &c = ctrls[mid];
    if (at_end) rec[mid].lam2 = c.maxlam[1];
    else { rec[mid].t = c.t; rec[mid].dt = c.dt; rec[mid].tide = c.tide; }
}

// nodal gradients: sum of the per-element-vertex adjoints over the elements
// around each node
__global__ __launch_bounds__(256) void k_adj_gather(MeshDev m, const double *__restrict__ zvb,
                                                    const double *__restrict__ nvb, double *__restrict__ gz,
                                                    double *__restrict__ gn)
{
    const int i = xcd_block() * 256 + (int)threadIdx.x;
    if (i >= m.nn) return;
    const int P = m.pitch;
    double sz = 0.0, sn = 0.0;
    const int b = m.n2e_off[i], en = m.n2e_off[i + 1];
    for (int j = b; j < en; ++j) {
        const int v = m.n2e[j];
        const int el = v & N2E_EL_MASK, k = (unsigned)v >> N2E_K_SHIFT;
        sz += zvb[k * P + el];
        sn += nvb[k * P + el];
    }
    gz[i] = sz;
    gn[i] = sn;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_adj_limit(const MeshDev &m, const PhysDev &ph, const double *U, const double *nodemm, const double *Sb,
                      const double *slotb, const double *zbe, double *zvb, double *Ub, double *vmb, AdjCtrl *actl,
                      const AdjScalarArgs *sa, hipStream_t s)
{
    const AdjScalarArgs none{};
    hipLaunchKernelGGL(k_adj_limit, dim3((unsigned)(m.pitch / 256)), dim3(256), 0, s, m, ph, U, nodemm, Sb, slotb, zbe,
                       zvb, Ub, vmb, actl, sa ? *sa : none);
}

void launch_adj_node_sums(const MeshDev &m, const double *vmb, double *nsum, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_node_sums, dim3((unsigned)((m.nn + 255) / 256)), dim3(256), 0, s, m, vmb, nsum);
}

void launch_adj_rhs(int mode, const MeshDev &m, const PhysDev &ph, const double *W, const double *B,
                    const double *U, const double *Ubown, const double *nsum, const int *nodearg, const int *wd,
                    const StepRec *rec, double *Inb, double *Bb, double *Fhb, double *zvb, double *nvb,
                    double *dtparts, hipStream_t s)
{
    const dim3 g((unsigned)(m.pitch / 256)), b(256);
    if (mode == 1)
        hipLaunchKernelGGL(k_adj_rhs<1>, g, b, 0, s, m, ph, W, B, U, Ubown, nsum, nodearg, wd, rec, Inb, Bb, Fhb,
                           zvb, nvb, dtparts);
    else
        hipLaunchKernelGGL(k_adj_rhs<2>, g, b, 0, s, m, ph, W, B, U, Ubown, nsum, nodearg, wd, rec, Inb, Bb, Fhb,
                           zvb, nvb, dtparts);
}

// the elements whose wet/dry flag is not wet, counted into *out
__global__ __launch_bounds__(256) void k_count_notwet(MeshDev m, const int *__restrict__ wd, int *out)
{
    const ElemLane L = elem_lane(m);
    const unsigned long long bal = __ballot(L.valid && wd[L.e] != 1);
    if (((int)threadIdx.x & 63) == 0 && bal) atomicAdd(out, __popcll(bal));
}

void launch_count_notwet(const MeshDev &m, const int *wd, int *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_count_notwet, dim3((unsigned)(m.pitch / 256)), dim3(256), 0, s, m, wd, out);
}

void launch_adj_edge(const MeshDev &m, const PhysDev &ph, const double *win, const double *bzeta, const double *bqn,
                     const int *wd, const double *Fhb, const StepRec *rec, const StepRec *rec_lam, AdjCtrl *actl,
                     double *slotb, double *zbe, double *tideparts, int with_lambda, const AdjEdgeQueue &eq,
                     hipStream_t s)
{
    const int nblk = m.ned_pad / 256;
    const dim3 g((unsigned)nblk), b(256);
    if (ph.wetdry && eq.split) {
        // the queue counter alternates between calls; each dry launch zeroes
        // the one the next call uses (both start at zero)
        int *qc = eq.qcount + eq.parity, *qn = eq.qcount + (1 - eq.parity);
        hipLaunchKernelGGL(k_adj_edge_wet, g, b, 0, s, m, ph, win, bzeta, bqn, wd, Fhb, rec, rec_lam, actl,
                           with_lambda, slotb, zbe, tideparts, eq.blist, eq.tb, eq.queue, qc);
        // two workgroups per CU at the dry kernel's 2 waves/SIMD
        const dim3 gd((unsigned)(nblk < 512 ? nblk : 512));
        hipLaunchKernelGGL(k_adj_edge_dry, gd, b, 0, s, m, ph, win, bzeta, bqn, wd, Fhb, rec, rec_lam, actl,
                           with_lambda, slotb, zbe, tideparts, eq.blist, eq.tb, eq.queue, qc, qn);
        eq.parity = 1 - eq.parity;
    } else if (ph.wetdry)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_edge<true>), g, b, 0, s, m, ph, win, bzeta, bqn, wd, Fhb, rec,
                           rec_lam, actl, with_lambda, slotb, zbe, tideparts);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_edge<false>), g, b, 0, s, m, ph, win, bzeta, bqn, wd, Fhb, rec,
                           rec_lam, actl, with_lambda, slotb, zbe, tideparts);
}

// the tie count of the recorded max |lambda| that k_adj_edge with
// with_lambda divides the max-|lambda| adjoint by (after k_adj_scalar)
void launch_adj_ties(const MeshDev &m, const PhysDev &ph, const double *win, const double *bzeta, const double *bqn,
                     const int *wd, const StepRec *rec, const StepRec *rec_lam, AdjCtrl *actl, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_ties, dim3((unsigned)(m.ned_pad / 256)), dim3(256), 0, s, m, ph, win, bzeta, bqn, wd,
                       rec_lam, rec, actl);
}

void launch_adj_combine(const MeshDev &m, double *Dst, const double *slotb, const double *zbe, double *zvb,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_combine, dim3((unsigned)(m.pitch / 256)), dim3(256), 0, s, m, Dst, slotb, zbe, zvb);
}

void launch_adj_scalar(const MeshDev &m, const AdjScalarArgs &sa, AdjCtrl *actl, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_scalar, dim3((unsigned)m.nmem), dim3(256), 0, s, m, sa, actl);
}

void launch_adj_gauges(const MeshDev &m, int ngauge, const int *gauge_el, const double *hval, const double *obs,
                       double *Wb, double *zvb, hipStream_t s)
{
    const int n = m.nmem * ngauge;
    hipLaunchKernelGGL(k_adj_gauges, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, ngauge, gauge_el, hval,
                       obs, Wb, zvb);
}

void launch_gauge_sample(const MeshDev &m, int ngauge, const int *gauge_el, const double *w, double *hval,
                         const double *obs, AdjCtrl *actl, hipStream_t s, int mstride)
{
    const int n = m.nmem * ngauge;
    if (mstride <= 0) mstride = ngauge;
    hipLaunchKernelGGL(k_gauge_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, ngauge, gauge_el, w,
                       hval, mstride);
    if (obs && actl)
        hipLaunchKernelGGL(k_gauge_loss, dim3((unsigned)((m.nmem + 63) / 64)), dim3(64), 0, s, m, ngauge, hval, obs,
                           actl);
}

void launch_step_record(const MeshDev &m, const Ctrl *ctrls, StepRec *rec, int at_end, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_record, dim3((unsigned)((m.nmem + 63) / 64)), dim3(64), 0, s, m, ctrls, rec, at_end);
}

void launch_adj_gather(const MeshDev &m, const double *zvb, const double *nvb, double *gz, double *gn, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_gather, dim3((unsigned)((m.nn + 255) / 256)), dim3(256), 0, s, m, zvb, nvb, gz, gn);
}
