// lbib_rem.hip -- lattice kernel for LBIB_FP16_REM storage: FP32 moments
// plus FP16 remainders (the layout and the encoding are described with
// rem_deq in lbib_device_math.h).
//
// Populations are stored as the moments of the post-collision state at the
// site that computed them, (rho - 1, jx, jy) for the fluid and C for each
// species, in FP32, plus each population's distance from the equilibrium of
// those moments (fluid) or from t_q C (species) in FP16.  Pull streaming
// rebuilds population q at site x from the moments and the remainder stored
// at x - e_q.  The remainders are the non-equilibrium parts, small against
// the populations, so their FP16 rounding costs far less than rounding the
// deviations f - w themselves (LBIB_FP16_DEV), while a site moves
// 30 + 14 ns bytes each way instead of 36 + 20 ns in FP32.
//
// The physics is that of the other mixed-precision kernels (FluidSolver::
// collide, FluidSolver.cpp:103-183 with Luo1993ForceModel.cpp:29-70 or
// GuoZhengShi2002ForceModel.cpp:29-78; CDESolverD2Q5::collide,
// CDESolverD2Q5.cpp:75-107 and the registered reactions; the no-flux
// boundary of BoundarySolverNoFluxD2Q5.cpp:37-113; MassSolverBoxOutlet.cpp:
// 42-72): only the loads and stores differ.
//
// Roofline: HBM bound; algorithmic bytes per site = 2 (30 + 14 ns).
#include "lbib_mixed.h"

namespace lbib {

template <int NS, int FM>
__global__ __launch_bounds__(256) void k_stream_collide_rem(const MixedArgs a) {
    if (step_stopped(a.st)) return;
    const int W = (int)blockDim.x;
    const uint32_t wg = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int x = (int)(wg % gridDim.x) * W + (int)threadIdx.x;
    const int y = (int)(wg / gridDim.x);
    if (x >= a.nx) return;
    Frame fr;
    if (!site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, x, y, fr)) return;
    constexpr int NSS = NS > 0 ? NS : 1;
    const long n = a.n;
    const long r0 = (long)y * a.nx;
    const int xm = (x == fr.x0) ? fr.x0 + fr.w - 1 : x - 1;
    const int xp = (x == fr.x0 + fr.w - 1) ? fr.x0 : x + 1;
    const long rS = (long)((y == fr.y0) ? fr.y0 + fr.h - 1 : y - 1) * a.nx;
    const long rN = (long)((y == fr.y0 + fr.h - 1) ? fr.y0 : y + 1) * a.nx;
    const long i = r0 + x;
    // source site of the population pulled along each direction
    const long src[9] = {i, r0 + xm, rS + x, r0 + xp, rN + x, rS + xm, rS + xp, rN + xp, rN + xm};
    const float* M = static_cast<const float*>(a.fs);
    const uint16_t* R = reinterpret_cast<const uint16_t*>(M + 3 * n);
    const float* Cg = static_cast<const float*>(a.gs);
    const uint16_t* Rg = reinterpret_cast<const uint16_t*>(Cg + (long)NSS * n);

    float mo[9][3];
    uint16_t rf[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        mo[q][0] = M[src[q]];
        mo[q][1] = M[n + src[q]];
        mo[q][2] = M[2 * n + src[q]];
        rf[q] = __builtin_nontemporal_load(&R[(long)q * n + src[q]]);   // read once per step
    }
    float cs[NSS][5];
    uint16_t rs[NSS][5];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) {
            cs[s][q] = Cg[(long)s * n + src[q]];
            rs[s][q] = __builtin_nontemporal_load(&Rg[(long)(5 * s + q) * n + src[q]]);
        }
    const uint32_t tg = a.tag[i];
    float Fx = 0.0f, Fy = 0.0f;
    if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (x >> 4)]) {
        const float2 F = a.F[i];
        Fx = F.x;
        Fy = F.y;
    }
    float d[9];
#pragma unroll
    for (int q = 0; q < 9; q++) d[q] = rem_deq(q, mo[q][0], mo[q][1], mo[q][2]) + rem_dec(rf[q]);
    float g[NSS][5];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) g[s][q] = rem_t(q) * cs[s][q] + rem_dec(rs[s][q]);

    const uint32_t dom = tg & DOM_MASK;
    float m = 0.0f;
    if (a.msite) { m = (float)a.msite[i]; a.msite[i] = 0.0; }
    else if (a.mtab) m = (dom < (uint32_t)a.ntab) ? (float)a.mtab[dom] : 0.0f;
    float pu = 0.0f, pv = 0.0f;
    if (FM != 0 || (NS > 0 && (tg >> BC_SHIFT))) {
        const float2 up = *reinterpret_cast<const float2*>(a.u + 2 * i);
        pu = up.x;
        pv = up.y;
    }
    if (NS > 0 && (tg >> BC_SHIFT)) {
        // no-flux cut links: incoming population from the first listed
        // species' pre-stream state, swapped into every listed species
        const int inv[5] = {DT, DW, DS, DE, DN};
        for (int o = 1; o <= 4; o++) {
            if (!((tg >> (BC_SHIFT + o - 1)) & 1u)) continue;
            const int dd = inv[o];
            const long e = (long)(o - 1) * n + i;
            const int desc = a.bc_desc[e];
            const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
            if (b1 <= b0) continue;
            const int s1 = a.bcd_sp[b0];
            const float c1 = Cg[(long)s1 * n + i];
            float gp[5];   // the pre-stream populations of the site
            for (int q = 0; q < 5; q++) gp[q] = rem_t(q) * c1 + rem_dec(Rg[(long)(5 * s1 + q) * n + i]);
            const float Cp = gp[0] + gp[1] + gp[2] + gp[3] + gp[4];
            const float c6 = Cp / 6.0f;
            const float bvx = (float)a.bc_v[2 * e], bvy = (float)a.bc_v[2 * e + 1];
            float feq, eq;
            if (dd == DE) { feq = c6 * (1.0f + 3.0f * bvx); eq = c6 * (1.0f + 3.0f * pu); }
            else if (dd == DN) { feq = c6 * (1.0f + 3.0f * bvy); eq = c6 * (1.0f + 3.0f * pv); }
            else if (dd == DW) { feq = c6 * (1.0f - 3.0f * bvx); eq = c6 * (1.0f - 3.0f * pu); }
            else { feq = c6 * (1.0f - 3.0f * bvy); eq = c6 * (1.0f - 3.0f * pv); }
            float carry = feq + (gp[dd] - eq);
            for (int kk = b0; kk < b1; kk++) {
                const int sk = a.bcd_sp[kk];
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (s == sk) { const float t = g[s][dd]; g[s][dd] = carry; carry = t; }
            }
        }
    }

    float uxn, uyn;
    const float tauI = fr.r < 0 ? a.tauI : (float)a.rects[fr.r].tauI;
    const float tau = fr.r < 0 ? a.tau : (float)a.rects[fr.r].tau;
    collide_dev_f32<FM>(d, m, Fx, Fy, tauI, tau, pu, pv, uxn, uyn);
    if (NS > 0) {
        float cj[NSS], cseq[NSS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            cj[s] = g[s][0] + g[s][1] + g[s][2] + g[s][3] + g[s][4];
            cseq[s] = cj[s];
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const float tI = fr.r < 0 ? a.sp[s].tauI : (float)a.rects[fr.r].tauIs[s];
            species_collide_f32<NS>(g[s], cj[s], uxn, uyn, a.sp[s], tI, dom, iter_of(a.it, a.st), cj, cseq);
            // a later species sees this one's concentration after its collide
            cseq[s] = g[s][0] + g[s][1] + g[s][2] + g[s][3] + g[s][4];
        }
    }
    if (a.box_outlet && (x == fr.x0 || y == fr.y0)) {
        const int passes = (x == fr.x0) + (y == fr.y0);
        for (int p = 0; p < passes; p++) {
            float sd = 0.0f;
#pragma unroll
            for (int q = 0; q < 9; q++) sd += d[q];
            const float sf = 1.0f / (1.0f + sd);
#pragma unroll
            for (int q = 0; q < 9; q++) d[q] = (d[q] - wq(q) * sd) * sf;   // (w + d) / (1 + sd) - w
#pragma unroll
            for (int s = 0; s < NS; s++)
#pragma unroll
                for (int q = 0; q < 5; q++) g[s][q] *= sf;
        }
    }

    float* Md = static_cast<float*>(a.fd);
    uint16_t* Rd = reinterpret_cast<uint16_t*>(Md + 3 * n);
    float rm1, jx, jy;
    rem_moments(d, rm1, jx, jy);
    __builtin_nontemporal_store(rm1, &Md[i]);
    __builtin_nontemporal_store(jx, &Md[n + i]);
    __builtin_nontemporal_store(jy, &Md[2 * n + i]);
#pragma unroll
    for (int q = 0; q < 9; q++)
        __builtin_nontemporal_store(rem_enc(d[q] - rem_deq(q, rm1, jx, jy)), &Rd[(long)q * n + i]);
    float* Cd = static_cast<float*>(a.gd);
    uint16_t* Rgd = reinterpret_cast<uint16_t*>(Cd + (long)NSS * n);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float c = rem_csum(g[s]);
        __builtin_nontemporal_store(c, &Cd[(long)s * n + i]);
#pragma unroll
        for (int q = 0; q < 5; q++)
            __builtin_nontemporal_store(rem_enc(g[s][q] - rem_t(q) * c), &Rgd[(long)(5 * s + q) * n + i]);
    }
    *reinterpret_cast<float2*>(a.u + 2 * i) = make_float2(uxn, uyn);
}

// The same update with each moment loaded once per lane: a lane loads the
// moments of its own column in the three rows it pulls from, and takes the
// columns x - 1 and x + 1 from the neighbouring lanes (wavefront shuffles).
// Lanes 0 and 63 of a wavefront, and lanes at a periodic frame's first or
// last column (whose neighbour column is the frame's other end), load that
// column themselves.  Same arithmetic on the same values as
// k_stream_collide_rem, so the results are identical.
template <int NS, int FM>
__global__ __launch_bounds__(256) void k_stream_collide_rem_sh(const MixedArgs a) {
    if (step_stopped(a.st)) return;
    const int W = (int)blockDim.x;
    const uint32_t wg = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int xr = (int)(wg % gridDim.x) * W + (int)threadIdx.x;
    const int y = (int)(wg / gridDim.x);
    const int lane = (int)threadIdx.x & 63;
    // lanes past the row end or outside every rectangle take part in the
    // shuffles with a valid address and store nothing
    const int x = min(xr, a.nx - 1);
    Frame fr;
    const bool live = xr < a.nx && site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, x, y, fr);
    if (!live) { fr.x0 = 0; fr.y0 = 0; fr.w = a.nx; fr.h = a.ny; fr.r = -1; }
    constexpr int NSS = NS > 0 ? NS : 1;
    const long n = a.n;
    const long r0 = (long)y * a.nx;
    const int xm = (x == fr.x0) ? fr.x0 + fr.w - 1 : x - 1;
    const int xp = (x == fr.x0 + fr.w - 1) ? fr.x0 : x + 1;
    const long rS = (long)((y == fr.y0) ? fr.y0 + fr.h - 1 : y - 1) * a.nx;
    const long rN = (long)((y == fr.y0 + fr.h - 1) ? fr.y0 : y + 1) * a.nx;
    const long i = r0 + x;
    const long src[9] = {i, r0 + xm, rS + x, r0 + xp, rN + x, rS + xm, rS + xp, rN + xp, rN + xm};
    const float* M = static_cast<const float*>(a.fs);
    const uint16_t* R = reinterpret_cast<const uint16_t*>(M + 3 * n);
    const float* Cg = static_cast<const float*>(a.gs);
    const uint16_t* Rg = reinterpret_cast<const uint16_t*>(Cg + (long)NSS * n);
    const bool w_need = lane == 0 || x == fr.x0;
    const bool e_need = lane == 63 || x == fr.x0 + fr.w - 1;

    // own column: rows S (y - 1), 0 (y), N (y + 1)
    const long rows[3] = {rS, r0, rN};
    float own[3][3], lft[3][3], rgt[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) own[r][k] = M[(long)k * n + rows[r] + x];
    uint16_t rf[9];
#pragma unroll
    for (int q = 0; q < 9; q++) rf[q] = R[(long)q * n + src[q]];
    float cown[NSS][3], cl[NSS], cr[NSS];
    uint16_t rs[NSS][5];
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
        for (int r = 0; r < 3; r++) cown[s][r] = Cg[(long)s * n + rows[r] + x];
#pragma unroll
        for (int q = 0; q < 5; q++) rs[s][q] = Rg[(long)(5 * s + q) * n + src[q]];
    }
    if (w_need) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++) lft[r][k] = M[(long)k * n + rows[r] + xm];
#pragma unroll
        for (int s = 0; s < NS; s++) cl[s] = Cg[(long)s * n + r0 + xm];
    }
    if (e_need) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++) rgt[r][k] = M[(long)k * n + rows[r] + xp];
#pragma unroll
        for (int s = 0; s < NS; s++) cr[s] = Cg[(long)s * n + r0 + xp];
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float up = __shfl_up(own[r][k], 1, 64), dn = __shfl_down(own[r][k], 1, 64);
            lft[r][k] = w_need ? lft[r][k] : up;
            rgt[r][k] = e_need ? rgt[r][k] : dn;
        }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float up = __shfl_up(cown[s][1], 1, 64), dn = __shfl_down(cown[s][1], 1, 64);
        cl[s] = w_need ? cl[s] : up;
        cr[s] = e_need ? cr[s] : dn;
    }
    if (!live) return;

    const uint32_t tg = a.tag[i];
    float Fx = 0.0f, Fy = 0.0f;
    if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (x >> 4)]) {
        const float2 F = a.F[i];
        Fx = F.x;
        Fy = F.y;
    }
    // source column (0 left, 1 own, 2 right) and row (0 S, 1 same, 2 N) of each direction
    auto mom = [&](int col, int row, int k) { return col == 0 ? lft[row][k] : (col == 1 ? own[row][k] : rgt[row][k]); };
    constexpr int SCOL[9] = {1, 0, 1, 2, 1, 0, 2, 2, 0};
    constexpr int SROW[9] = {1, 1, 0, 1, 2, 0, 0, 2, 2};
    float d[9];
#pragma unroll
    for (int q = 0; q < 9; q++)
        d[q] = rem_deq(q, mom(SCOL[q], SROW[q], 0), mom(SCOL[q], SROW[q], 1), mom(SCOL[q], SROW[q], 2)) +
               rem_dec(rf[q]);
    float g[NSS][5];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float cq[5] = {cown[s][1], cl[s], cown[s][0], cr[s], cown[s][2]};
#pragma unroll
        for (int q = 0; q < 5; q++) g[s][q] = rem_t(q) * cq[q] + rem_dec(rs[s][q]);
    }
    const uint32_t dom = tg & DOM_MASK;
    float m = 0.0f;
    if (a.msite) { m = (float)a.msite[i]; a.msite[i] = 0.0; }
    else if (a.mtab) m = (dom < (uint32_t)a.ntab) ? (float)a.mtab[dom] : 0.0f;
    float pu = 0.0f, pv = 0.0f;
    if (FM != 0 || (NS > 0 && (tg >> BC_SHIFT))) {
        const float2 up = *reinterpret_cast<const float2*>(a.u + 2 * i);
        pu = up.x;
        pv = up.y;
    }
    if (NS > 0 && (tg >> BC_SHIFT)) {
        // no-flux cut links: incoming population from the first listed
        // species' pre-stream state, swapped into every listed species
        const int inv[5] = {DT, DW, DS, DE, DN};
        for (int o = 1; o <= 4; o++) {
            if (!((tg >> (BC_SHIFT + o - 1)) & 1u)) continue;
            const int dd = inv[o];
            const long e = (long)(o - 1) * n + i;
            const int desc = a.bc_desc[e];
            const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
            if (b1 <= b0) continue;
            const int s1 = a.bcd_sp[b0];
            const float c1 = Cg[(long)s1 * n + i];
            float gp[5];   // the pre-stream populations of the site
            for (int q = 0; q < 5; q++) gp[q] = rem_t(q) * c1 + rem_dec(Rg[(long)(5 * s1 + q) * n + i]);
            const float Cp = gp[0] + gp[1] + gp[2] + gp[3] + gp[4];
            const float c6 = Cp / 6.0f;
            const float bvx = (float)a.bc_v[2 * e], bvy = (float)a.bc_v[2 * e + 1];
            float feq, eq;
            if (dd == DE) { feq = c6 * (1.0f + 3.0f * bvx); eq = c6 * (1.0f + 3.0f * pu); }
            else if (dd == DN) { feq = c6 * (1.0f + 3.0f * bvy); eq = c6 * (1.0f + 3.0f * pv); }
            else if (dd == DW) { feq = c6 * (1.0f - 3.0f * bvx); eq = c6 * (1.0f - 3.0f * pu); }
            else { feq = c6 * (1.0f - 3.0f * bvy); eq = c6 * (1.0f - 3.0f * pv); }
            float carry = feq + (gp[dd] - eq);
            for (int kk = b0; kk < b1; kk++) {
                const int sk = a.bcd_sp[kk];
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (s == sk) { const float t = g[s][dd]; g[s][dd] = carry; carry = t; }
            }
        }
    }

    float uxn, uyn;
    const float tauI = fr.r < 0 ? a.tauI : (float)a.rects[fr.r].tauI;
    const float tau = fr.r < 0 ? a.tau : (float)a.rects[fr.r].tau;
    collide_dev_f32<FM>(d, m, Fx, Fy, tauI, tau, pu, pv, uxn, uyn);
    if (NS > 0) {
        float cj[NSS], cseq[NSS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            cj[s] = g[s][0] + g[s][1] + g[s][2] + g[s][3] + g[s][4];
            cseq[s] = cj[s];
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const float tI = fr.r < 0 ? a.sp[s].tauI : (float)a.rects[fr.r].tauIs[s];
            species_collide_f32<NS>(g[s], cj[s], uxn, uyn, a.sp[s], tI, dom, iter_of(a.it, a.st), cj, cseq);
            // a later species sees this one's concentration after its collide
            cseq[s] = g[s][0] + g[s][1] + g[s][2] + g[s][3] + g[s][4];
        }
    }
    if (a.box_outlet && (x == fr.x0 || y == fr.y0)) {
        const int passes = (x == fr.x0) + (y == fr.y0);
        for (int p = 0; p < passes; p++) {
            float sd = 0.0f;
#pragma unroll
            for (int q = 0; q < 9; q++) sd += d[q];
            const float sf = 1.0f / (1.0f + sd);
#pragma unroll
            for (int q = 0; q < 9; q++) d[q] = (d[q] - wq(q) * sd) * sf;   // (w + d) / (1 + sd) - w
#pragma unroll
            for (int s = 0; s < NS; s++)
#pragma unroll
                for (int q = 0; q < 5; q++) g[s][q] *= sf;
        }
    }

    float* Md = static_cast<float*>(a.fd);
    uint16_t* Rd = reinterpret_cast<uint16_t*>(Md + 3 * n);
    float rm1, jx, jy;
    rem_moments(d, rm1, jx, jy);
    __builtin_nontemporal_store(rm1, &Md[i]);
    __builtin_nontemporal_store(jx, &Md[n + i]);
    __builtin_nontemporal_store(jy, &Md[2 * n + i]);
#pragma unroll
    for (int q = 0; q < 9; q++)
        __builtin_nontemporal_store(rem_enc(d[q] - rem_deq(q, rm1, jx, jy)), &Rd[(long)q * n + i]);
    float* Cd = static_cast<float*>(a.gd);
    uint16_t* Rgd = reinterpret_cast<uint16_t*>(Cd + (long)NSS * n);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float c = rem_csum(g[s]);
        __builtin_nontemporal_store(c, &Cd[(long)s * n + i]);
#pragma unroll
        for (int q = 0; q < 5; q++)
            __builtin_nontemporal_store(rem_enc(g[s][q] - rem_t(q) * c), &Rgd[(long)(5 * s + q) * n + i]);
    }
    *reinterpret_cast<float2*>(a.u + 2 * i) = make_float2(uxn, uyn);
}

// FP64 natural-order populations of one field <-> the moments-and-remainders
// store.  field 0: the fluid (9 planes); field 1 + s: species s (5 planes).
__global__ void k_rem_pack(const double* __restrict__ src, void* dst, long n, int field, int nsp) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (field == 0) {
            float* M = static_cast<float*>(dst);
            uint16_t* R = reinterpret_cast<uint16_t*>(M + 3 * n);
            float d[9];
            for (int q = 0; q < 9; q++) d[q] = (float)(src[(long)q * n + i] - (double)rem_w(q));
            float rm1, jx, jy;
            rem_moments(d, rm1, jx, jy);
            M[i] = rm1;
            M[n + i] = jx;
            M[2 * n + i] = jy;
            for (int q = 0; q < 9; q++) R[(long)q * n + i] = rem_enc(d[q] - rem_deq(q, rm1, jx, jy));
        } else {
            const int s = field - 1;
            float* C = static_cast<float*>(dst);
            uint16_t* R = reinterpret_cast<uint16_t*>(C + (long)nsp * n);
            float gf[5];
            for (int q = 0; q < 5; q++) gf[q] = (float)src[(long)q * n + i];
            const float c = rem_csum(gf);
            C[(long)s * n + i] = c;
            for (int q = 0; q < 5; q++) R[(long)(5 * s + q) * n + i] = rem_enc(gf[q] - rem_t(q) * c);
        }
    }
}
__global__ void k_rem_unpack(const void* srcv, double* __restrict__ dst, long n, int field, int nsp) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (field == 0) {
            const float* M = static_cast<const float*>(srcv);
            const uint16_t* R = reinterpret_cast<const uint16_t*>(M + 3 * n);
            const float rm1 = M[i], jx = M[n + i], jy = M[2 * n + i];
            for (int q = 0; q < 9; q++)
                dst[(long)q * n + i] =
                    (double)(rem_deq(q, rm1, jx, jy) + rem_dec(R[(long)q * n + i])) + (double)rem_w(q);
        } else {
            const int s = field - 1;
            const float* C = static_cast<const float*>(srcv);
            const uint16_t* R = reinterpret_cast<const uint16_t*>(C + (long)nsp * n);
            const float c = C[(long)s * n + i];
            for (int q = 0; q < 5; q++)
                dst[(long)q * n + i] = (double)(rem_t(q) * c + rem_dec(R[(long)(5 * s + q) * n + i]));
        }
    }
}

static int rem_grid(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

int launch_rem_pack(lbib_ctx* c, const double* src_dev, int field) {
    void* dst = field == 0 ? c->f[c->cur] : c->g[c->cur];
    k_rem_pack<<<rem_grid(c->n), 256, 0, c->stream>>>(src_dev, dst, c->n, field, std::max(c->ns, 1));
    LBIB_HIP(hipGetLastError());
    return 0;
}
int launch_rem_unpack(lbib_ctx* c, double* dst_dev, int field) {
    const void* src = field == 0 ? c->f[c->cur] : c->g[c->cur];
    k_rem_unpack<<<rem_grid(c->n), 256, 0, c->stream>>>(src, dst_dev, c->n, field, std::max(c->ns, 1));
    LBIB_HIP(hipGetLastError());
    return 0;
}

template <int NS, int FM>
static void launch_rem_ns(lbib_ctx* c, const MixedArgs& a, int variant, int wg) {
    const dim3 grid((c->nx + wg - 1) / wg, c->ny);
    if (variant == 2) k_stream_collide_rem_sh<NS, FM><<<grid, wg, 0, c->stream>>>(a);
    else k_stream_collide_rem<NS, FM><<<grid, wg, 0, c->stream>>>(a);
}
template <int FM>
static void launch_rem_fm(lbib_ctx* c, const MixedArgs& a, int variant, int wg) {
    switch (c->ns) {
        case 0: launch_rem_ns<0, FM>(c, a, variant, wg); break;
        case 1: launch_rem_ns<1, FM>(c, a, variant, wg); break;
        case 2: launch_rem_ns<2, FM>(c, a, variant, wg); break;
        case 3: launch_rem_ns<3, FM>(c, a, variant, wg); break;
        default: launch_rem_ns<4, FM>(c, a, variant, wg); break;
    }
}

// variant 0: one load per pulled moment; variant 2: moments shared across
// lanes by shuffles
int launch_rem_shape(lbib_ctx* c, const MixedArgs& a, int variant, int wg) {
    if (c->cfg.force_model == LBIB_FORCE_LUO) launch_rem_fm<0>(c, a, variant, wg);
    else if (c->cfg.force_model == LBIB_FORCE_GUO_AS_WRITTEN) launch_rem_fm<1>(c, a, variant, wg);
    else launch_rem_fm<2>(c, a, variant, wg);
    LBIB_HIP(hipGetLastError());
    return 0;
}

}  // namespace lbib
