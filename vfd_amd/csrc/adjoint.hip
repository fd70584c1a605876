// adjoint.hip - reverse of a recorded DFSPH step (hand-written adjoint; no reference counterpart).
//
// Differentiated map: one step (x, v; mu) -> (x', v') with the step's neighbour lists, time steps, Jacobi iteration
// counts and clamp masks held fixed. Kernel values are the forward's table values; their derivatives are those of
// the cubic spline the table samples (Kernel/DFSPHKernels.h:86-131): dW/drho and d(gradW/r)/drho. The Jacobi solves
// are reversed iteration by iteration (their transposed iteration: the same neighbour sums, in reverse order); the
// viscosity solve by one transposed linear solve (BiCGSTAB). Lists are symmetric (no particle at the cap, checked
// when recording), so every position cotangent is gathered per particle: pair (i, j) contributes to both ends from
// i's list and from j's, without atomics.
//
// Boundary contact enters through the sample terms vb W(x - xb) (density) and vb gradW(x - xb) (factor, divergence,
// pressure acceleration), whose position derivatives the recording kernel (k_adj_boundary_tape) stores per particle
// and body from the volume map's interpolant (gradient of vb, Jacobian of r = x - xb including the normal's
// derivative), and through the push-out's Jacobian; a pushed particle's entry velocity has no effect (it is zeroed).
// With a boundary viscosity each contact's four tangential samples add a symmetric 3x3 block to its viscosity row;
// the transposed solve applies it, and its derivative follows the samples' offsets, the turning of their axes with
// the normal (the rows of dn/dx are recorded too), the sample volume and the density. Surface tension enters through
// its smoothing pass, with the recorded Monte-Carlo normals and curvatures held.
//
// Per-step arrays are in that step's sorted slot order; the host maps cotangents through the recorded ids.
#include "device.h"
#include "kernels.h"

namespace vfd {

struct AdjK {
    const float2* __restrict__ tab;
    float R, invStep, K, L, V, rho0, m, mu, R2;
    uint32_t n, cap;                    // particles of the scene being reversed, list capacity
    uint32_t base, stride;              // the scene's first slot; slots per record of the step (contact records)
    const uint32_t* __restrict__ count;
    const uint32_t* __restrict__ list;  // [j * cap + i]
    const float4* __restrict__ x;
    const float4* __restrict__ con;  // contact records [(kContactRecs b + q) stride + i] (k_adj_boundary_tape)
    uint32_t bodies;                 // 0: no contact records
    float muB, td;                   // boundary viscosity (0: its samples are not differentiated) and sample distance
    float invR, dG0, LR;             // 1 / R, 3 L / R^3 (d(gradW/r)/drho for q <= 1/2), L / R
};

__device__ __forceinline__ float adj_g(const AdjK& a, float rl) {  // gradW / r as the forward's table gives it
    if (!(rl <= a.R)) return 0.0f;
    const unsigned p0 = (unsigned)(rl * a.invStep);
    return a.tab[p0 < kTableEntries - 1u ? p0 : kTableEntries - 1u].y;
}
__device__ __forceinline__ float adj_w(const AdjK& a, float r2) {  // W as the forward's table gives it
    if (!(r2 <= a.R2)) return 0.0f;
    const unsigned p0 = (unsigned)(sqrtf(r2) * a.invStep);
    return a.tab[p0 < kTableEntries - 1u ? p0 : kTableEntries - 1u].x;
}
__device__ __forceinline__ float adj_dW(const AdjK& a, float rl) {  // spline dW/drho
    const float q = rl / a.R;
    if (!(q <= 1.0f)) return 0.0f;
    if (q <= 0.5f) return a.K * (18.0f * q * q - 12.0f * q) / a.R;
    return -6.0f * a.K * (1.0f - q) * (1.0f - q) / a.R;
}
// The pair terms of the reverse loops divide through hardware reciprocals (v_rcp_f32, about 1 ulp): the adjoint is an
// FP32 derivative held to a relative tolerance, not a bit-exact restatement, and IEEE divisions would dominate its
// arithmetic.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// spline d(gradW/r)/drho = (L/R) f (2 / (R rl) + f / rl^2), f = 1 - rl/R, for rl > R/2; ir = 1 / rl
__device__ __forceinline__ float adj_dG(const AdjK& a, float rl, float ir) {
    const float q = rl * a.invR;
    if (!(q <= 1.0f) || !(rl > 1.0e-5f)) return 0.0f;
    if (q <= 0.5f) return a.dG0;
    const float f = 1.0f - q;
    return a.LR * f * ir * (2.0f * a.invR + f * ir);
}
__device__ __forceinline__ float adj_dG(const AdjK& a, float rl) { return adj_dG(a, rl, rcp(rl)); }
// d(c . G(|r|) r)/dr = G c + G'(|r|) (c . r) / |r| r
__device__ __forceinline__ f3 vjp_grad(const AdjK& a, f3 r, float rl, float g, f3 c) {
    f3 o = g * c;
    if (rl > 0.0f) {
        const float ir = rcp(rl);
        o = o + (adj_dG(a, rl, ir) * dot3(c, r) * ir) * r;
    }
    return o;
}
__device__ __forceinline__ uint32_t adj_nbr(const AdjK& a, uint32_t i, uint32_t j) {
    return a.list[(size_t)j * a.cap + i];
}
__device__ __forceinline__ uint32_t pid() { return blockIdx.x * blockDim.x + threadIdx.x; }
// The neighbour-loop kernels run over the scene's slots [base, base + n) (one thread each) and index every array by
// slot, as the lists do; the pointwise kernels take arrays offset to the scene's first slot.
__device__ __forceinline__ bool adj_slot(const AdjK& a, uint32_t& i) {
    i = a.base + pid();
    return pid() < a.n;
}
// Boundary sample records of particle i, body b: {r = x - xb, vb}, {grad vb}, rows of dr/dx.
struct AdjContact {
    f3 r, dv, j0, j1, j2;
    float vb;
};
__device__ __forceinline__ bool adj_contact(const AdjK& a, uint32_t b, uint32_t i, AdjContact& c) {
    const float4* p = a.con + (size_t)kContactRecs * b * a.stride + i;
    const float4 r = p[0];
    if (!(r.w > 0.0f)) return false;
    c.r = mk3(r.x, r.y, r.z);
    c.vb = r.w;
    c.dv = ld3(p[a.stride]);
    c.j0 = ld3(p[2 * (size_t)a.stride]);
    c.j1 = ld3(p[3 * (size_t)a.stride]);
    c.j2 = ld3(p[4 * (size_t)a.stride]);
    return true;
}
// sum over bodies of vb gradW(r): the boundary part of particle i's kernel-gradient sum
__device__ __forceinline__ f3 adj_contact_sum(const AdjK& a, uint32_t i) {
    f3 s = mk3(0.0f, 0.0f, 0.0f);
    AdjContact c;
    for (uint32_t b = 0; b < a.bodies; ++b)
        if (adj_contact(a, b, i, c)) s = s + (c.vb * adj_g(a, sqrtf(dot3(c.r, c.r)))) * c.r;
    return s;
}
// d/dx_i of q . adj_contact_sum(i) for a fixed q: grad vb (gradW . q) + (dr/dx)^T vjp(r, vb q)
__device__ __forceinline__ f3 adj_contact_vjp(const AdjK& a, uint32_t i, f3 q) {
    f3 o = mk3(0.0f, 0.0f, 0.0f);
    AdjContact c;
    for (uint32_t b = 0; b < a.bodies; ++b) {
        if (!adj_contact(a, b, i, c)) continue;
        const float rl = sqrtf(dot3(c.r, c.r)), g = adj_g(a, rl);
        const f3 u = vjp_grad(a, c.r, rl, g, c.vb * q);
        o = o + (g * dot3(c.r, q)) * c.dv + (u.x * c.j0 + u.y * c.j1 + u.z * c.j2);
    }
    return o;
}
// d/dx_i of sum over bodies of vb W(r)
__device__ __forceinline__ f3 adj_contact_density_grad(const AdjK& a, uint32_t i) {
    f3 o = mk3(0.0f, 0.0f, 0.0f);
    AdjContact c;
    for (uint32_t b = 0; b < a.bodies; ++b) {
        if (!adj_contact(a, b, i, c)) continue;
        const float r2 = dot3(c.r, c.r), rl = sqrtf(r2);
        const float w = adj_w(a, r2), dw = rl > 0.0f ? c.vb * adj_dW(a, rl) / rl : 0.0f;
        o = o + w * c.dv + (dw * c.r.x * c.j0 + dw * c.r.y * c.j1 + dw * c.r.z * c.j2);
    }
    return o;
}

// Boundary viscosity samples of a contact (tangential(), DFSPHKernels.cu:636-681, 793-839, formed from the recorded
// r = x - xb as the forward forms them): N = -r / |r|, v the axis GetOrthogonalVectors picks (held), x = N x v,
// y = N x x, t1 = x / |x|, t2 = y / |y|; the offsets are pd = r + td t1, r - td t1, r + td t2, r - td t2.
struct AdjTan {
    f3 N, v, x, t1, t2;
    float xl, yl;
};
__device__ __forceinline__ bool adj_tan(f3 r, AdjTan& t) {
    const f3 n0 = -r;
    const float nl = sqrtf(dot3(n0, n0));
    if (!(nl > 0.0001f)) return false;
    t.N = n0 / nl;
    t.v = mk3(1.0f, 0.0f, 0.0f);
    if ((double)fabsf(dot3(t.v, t.N)) > 0.999) t.v = mk3(0.0f, 1.0f, 0.0f);
    t.x = cross(t.N, t.v);
    const f3 y = cross(t.N, t.x);
    t.xl = sqrtf(dot3(t.x, t.x));
    t.yl = sqrtf(dot3(y, y));
    t.t1 = t.x / t.xl;
    t.t2 = y / t.yl;
    return true;
}
__device__ __forceinline__ f3 adj_tan_pd(const AdjContact& c, const AdjTan& t, float td, int k) {
    const f3 tm = k < 2 ? t.t1 : t.t2;
    return (k & 1) ? c.r - td * tm : c.r + td * tm;
}
// C_i w = sum over the particle's contacts and their four samples of c G(|pd|) pd (pd . w),
// c = 10 mu_b (vb / 4) / (|pd|^2 + 0.01 R^2): the boundary block of the viscosity row (symmetric)
__device__ __forceinline__ f3 adj_visc_boundary(const AdjK& a, uint32_t i, f3 w) {
    f3 o = mk3(0.0f, 0.0f, 0.0f);
    AdjContact c;
    for (uint32_t b = 0; b < a.bodies; ++b) {
        if (!adj_contact(a, b, i, c)) continue;
        AdjTan t;
        if (!adj_tan(c.r, t)) continue;
        const float vol = 0.25f * c.vb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f3 pd = adj_tan_pd(c, t, a.td, k);
            const float pl = sqrtf(dot3(pd, pd)), q = pl * pl + 0.01f * a.R2;
            o = o + (10.0f * a.muB * vol / q * adj_g(a, pl) * dot3(pd, w)) * pd;
        }
    }
    return o;
}
// d/dx_i of lb . C_i(x_i) y and its value: the samples move with the contact's offset r (dr/dx) and turn with its
// normal (dn/dx, through t1 and t2), and scale with vb
__device__ __forceinline__ f3 adj_visc_boundary_vjp(const AdjK& a, uint32_t i, f3 lb, f3 y) {
    f3 o = mk3(0.0f, 0.0f, 0.0f);
    AdjContact c;
    for (uint32_t b = 0; b < a.bodies; ++b) {
        if (!adj_contact(a, b, i, c)) continue;
        AdjTan t;
        if (!adj_tan(c.r, t)) continue;
        const float4* p = a.con + (size_t)kContactRecs * b * a.stride + i;
        const f3 n0 = ld3(p[5 * (size_t)a.stride]), n1 = ld3(p[6 * (size_t)a.stride]), n2 = ld3(p[7 * (size_t)a.stride]);
        const float vol = 0.25f * c.vb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f3 pd = adj_tan_pd(c, t, a.td, k);
            const float pl = sqrtf(dot3(pd, pd)), q = pl * pl + 0.01f * a.R2;
            const float ck = 10.0f * a.muB * vol / q, g = adj_g(a, pl);
            const float lp = dot3(lb, pd), py = dot3(pd, y);
            const float fk = ck * g * lp * py;
            f3 gp = (ck * g) * (py * lb + lp * y) + (lp * py * g * (-2.0f * ck / q)) * pd;
            if (pl > 0.0f) gp = gp + (lp * py * ck * adj_dG(a, pl) / pl) * pd;
            // (d pd / dx)^T gp = (dr/dx)^T gp -+ td (dn/dx)^T (dt/dN)^T gp, dN = -dn
            f3 w;
            if (k < 2) {
                const f3 u = (gp - dot3(t.t1, gp) * t.t1) / t.xl;
                w = cross(t.v, u);
            } else {
                const f3 u = (gp - dot3(t.t2, gp) * t.t2) / t.yl;
                w = cross(t.x, u) - cross(t.v, cross(t.N, u));
            }
            const f3 jn = w.x * n0 + w.y * n1 + w.z * n2;
            const f3 jr = gp.x * c.j0 + gp.y * c.j1 + gp.z * c.j2;
            o = o + ((k & 1) ? jr + a.td * jn : jr - a.td * jn) + (fk / c.vb) * c.dv;
        }
    }
    return o;
}

// The reverse neighbour loops fetch kAdjBatch list entries, then those neighbours' data, before computing any of
// them, so that the dependent loads overlap; the terms are still added in list order.
constexpr uint32_t kAdjBatch = 4;

// lk_i -= V sum_j (la_i - la_j) . gradW_ij + la_i . B_i          (A^T la = -D(la))
// lx_i += sum_j vjp(r_ij, -V (k_i + k_j)(la_i - la_j)) + d/dx_i (-k_i la_i . B_i)  (positions through A(k))
// with B_i = sum_b vb gradW(x_i - xb) the boundary samples' term. term() takes the neighbours in list order with
// g = gradW / r of the pair as the forward's table gives it.
struct AdjAccelT {
    f3 xi, lai, gx;
    float ki, d;
    __device__ __forceinline__ AdjAccelT(const AdjK& a, uint32_t i, const float* __restrict__ kap,
                                         const float4* __restrict__ la)
        : xi(ld3(a.x[i])), lai(ld3(la[i])), gx(mk3(0.0f, 0.0f, 0.0f)), ki(kap[i]), d(0.0f) {}
    __device__ __forceinline__ void term(const AdjK& a, f3 xj, f3 laj, float kj, float g) {
        const f3 r = xi - xj;
        const float rl = sqrtf(dot3(r, r));
        const f3 dl = lai - laj;
        d += dot3(dl, g * r);
        gx = gx + vjp_grad(a, r, rl, g, (-a.V * (ki + kj)) * dl);
    }
    __device__ __forceinline__ void finish(const AdjK& a, uint32_t i, float* lk, float4* lx) {
        if (a.bodies) {
            d = a.V * d + dot3(lai, adj_contact_sum(a, i));
            gx = gx + adj_contact_vjp(a, i, (-ki) * lai);
        } else {
            d = a.V * d;
        }
        lk[i] -= d;
        lx[i] = st4(ld3(lx[i]) + gx);
    }
};
__global__ void __launch_bounds__(kTile) k_adj_accel_T(AdjK a, const float* __restrict__ kap,
                                                       const float4* __restrict__ la, float* lk, float4* lx) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    AdjAccelT t(a, i, kap, la);
    const uint32_t cnt = a.count[i];
    for (uint32_t j0 = 0; j0 < cnt; j0 += kAdjBatch) {
        f3 xj[kAdjBatch], laj[kAdjBatch];
        float kj[kAdjBatch];
#pragma unroll
        for (uint32_t u = 0; u < kAdjBatch; ++u) {
            const uint32_t j = adj_nbr(a, i, min(j0 + u, cnt - 1u));
            xj[u] = ld3(a.x[j]);
            laj[u] = ld3(la[j]);
            kj[u] = kap[j];
        }
#pragma unroll
        for (uint32_t u = 0; u < kAdjBatch; ++u) {
            if (j0 + u >= cnt) break;
            const f3 r = t.xi - xj[u];
            t.term(a, xj[u], laj[u], kj[u], adj_g(a, sqrtf(dot3(r, r))));
        }
    }
    t.finish(a, i, lk, lx);
}

// lu_i += V sum_j (lf_i + lf_j) gradW_ij + lf_i B_i                       (D^T lf = -A(lf))
// lx_i += sum_j vjp(r_ij, V (lf_i + lf_j)(u_i - u_j)) + d/dx_i (lf_i u_i . B_i)  (positions through D(u))
struct AdjDivT {
    f3 xi, ui, s, gx;
    float lfi;
    __device__ __forceinline__ AdjDivT(const AdjK& a, uint32_t i, const float* __restrict__ lf,
                                       const float4* __restrict__ u)
        : xi(ld3(a.x[i])), ui(ld3(u[i])), s(mk3(0.0f, 0.0f, 0.0f)), gx(mk3(0.0f, 0.0f, 0.0f)), lfi(lf[i]) {}
    __device__ __forceinline__ void term(const AdjK& a, f3 xj, f3 uj, float lfj, float g) {
        const f3 r = xi - xj;
        const float rl = sqrtf(dot3(r, r));
        const float l2 = lfi + lfj;
        s = s + (l2 * g) * r;
        gx = gx + vjp_grad(a, r, rl, g, (a.V * l2) * (ui - uj));
    }
    __device__ __forceinline__ void finish(const AdjK& a, uint32_t i, float4* lu, float4* lx) {
        s = a.V * s;
        if (a.bodies) {
            s = s + lfi * adj_contact_sum(a, i);
            gx = gx + adj_contact_vjp(a, i, lfi * ui);
        }
        lu[i] = st4(ld3(lu[i]) + s);
        lx[i] = st4(ld3(lx[i]) + gx);
    }
};
__global__ void __launch_bounds__(kTile) k_adj_div_T(AdjK a, const float* __restrict__ lf,
                                                     const float4* __restrict__ u, float4* lu, float4* lx) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    AdjDivT t(a, i, lf, u);
    const uint32_t cnt = a.count[i];
    for (uint32_t j0 = 0; j0 < cnt; j0 += kAdjBatch) {
        f3 xj[kAdjBatch], uj[kAdjBatch];
        float lfj[kAdjBatch];
#pragma unroll
        for (uint32_t q = 0; q < kAdjBatch; ++q) {
            const uint32_t j = adj_nbr(a, i, min(j0 + q, cnt - 1u));
            xj[q] = ld3(a.x[j]);
            uj[q] = ld3(u[j]);
            lfj[q] = lf[j];
        }
#pragma unroll
        for (uint32_t q = 0; q < kAdjBatch; ++q) {
            if (j0 + q >= cnt) break;
            const f3 r = t.xi - xj[q];
            t.term(a, xj[q], uj[q], lfj[q], adj_g(a, sqrtf(dot3(r, r))));
        }
    }
    t.finish(a, i, lu, lx);
}

// The same two loops per neighbour group (the forward's group loop): the recorded step's halos, rebuilt from its lists
// by k_build_halo, are staged in LDS with records {x_j, s_j} {v_j} and the kernel table, and each pair reads its
// table entry through the code k_build_halo stored (the entry adj_g reads); a group whose halo does not fit gathers
// from the global arrays. Terms and their order are those of the per-particle kernels, so both give the same bits.
struct AdjStage {
    TileGrid tg;
    HaloView H;
    const uint32_t* __restrict__ local;  // [j * cap + i] halo slot | table code << 16
};
constexpr size_t kLdsAdjStage = kLdsG + HaloRec<LoadXSV>::kBytes;
template <class Make, class Fin>
__device__ __forceinline__ void adj_group_loop(const AdjK& a, const AdjStage& g, const float* __restrict__ s,
                                               const float4* __restrict__ v, Make&& make, Fin&& fin) {
    extern __shared__ float4 lds[];
    float* sG = reinterpret_cast<float*>(lds);
    load_table_g(sG, a.tab);
    float* rec = reinterpret_cast<float*>(lds + kLdsG / sizeof(float4));
    const LoadXSV load{a.x, s, v};
    const NbrSrc global{nullptr, g.local, a.list, a.count, a.cap};
    group_loop<2>(g.tg, g.H, rec, global, load, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const uint32_t i = tile * kTile + (threadIdx.x & (kTile - 1));
        if (!ok || i - a.base >= a.n) return;  // (unsigned: slots below base wrap past n)
        auto t = make(i);
        each_neighbor_rec<2>(src, i, a.R, a.invStep, t.xi, sG, load,
                             [&](float4 lo, float4 hi, float gv) { t.term(a, ld3(lo), ld3(hi), lo.w, gv); });
        fin(t, i);
    });
}
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_adj_accel_T_groups(AdjK a, AdjStage g,
                                                                                 const float* __restrict__ kap,
                                                                                 const float4* __restrict__ la,
                                                                                 float* lk, float4* lx) {
    adj_group_loop(
        a, g, kap, la, [&](uint32_t i) { return AdjAccelT(a, i, kap, la); },
        [&](AdjAccelT& t, uint32_t i) { t.finish(a, i, lk, lx); });
}
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_adj_div_T_groups(AdjK a, AdjStage g,
                                                                               const float* __restrict__ lf,
                                                                               const float4* __restrict__ u,
                                                                               float4* lu, float4* lx) {
    adj_group_loop(
        a, g, lf, u, [&](uint32_t i) { return AdjDivT(a, i, lf, u); },
        [&](AdjDivT& t, uint32_t i) { t.finish(a, i, lu, lx); });
}

// one reversed Jacobi iteration, pointwise part: kappa^{k+1} = kappa^k - s r^k, r^k = M (src - c f^k)
//   ls += -r lk, lsrc += M lr, lf = -c M lr, lr = -s lk, s = alpha / div
__global__ void __launch_bounds__(kTile) k_adj_jacobi_point(uint32_t n, const float* __restrict__ lk,
                                                            const float* __restrict__ alpha, float div,
                                                            const float* __restrict__ res, float c, float* lsrc,
                                                            float* ls, float* lf) {
    const uint32_t i = pid();
    if (i >= n) return;
    const float s = alpha[i] / div, r = res[i], l = lk[i];
    const bool m = r < 0.0f;
    const float lr = -s * l;
    ls[i] += -r * l;
    lsrc[i] += m ? lr : 0.0f;
    lf[i] = m ? -c * lr : 0.0f;
}

// pressure initialisation reversed: kp0 = mp0 (rhoadv - 1) gamma, rhoadv = rho / rho0 + H D(v2), gamma = alpha / H^2
__global__ void __launch_bounds__(kTile) k_adj_pres_init(uint32_t n, float H, float rho0,
                                                         const float* __restrict__ alpha,
                                                         const float* __restrict__ rhoadv, const float* __restrict__ lk,
                                                         const float* __restrict__ lsrc, const float* __restrict__ lgam,
                                                         float* lalpha, float* lrho, float* ld2) {
    const uint32_t i = pid();
    if (i >= n) return;
    const float gam = alpha[i] / (H * H), ra = rhoadv[i];
    const bool m = 1.0f - ra < 0.0f;
    const float lra = -lsrc[i] + (m ? gam * lk[i] : 0.0f);
    const float lg = lgam[i] + (m ? (ra - 1.0f) * lk[i] : 0.0f);
    lalpha[i] += lg / (H * H);
    lrho[i] += lra / rho0;
    ld2[i] = H * lra;
}

// divergence initialisation reversed: kv0 = D beta, D = max(D(v0), 0) (0 below 20 neighbours), beta = alpha / h
__global__ void __launch_bounds__(kTile) k_adj_div_init(uint32_t n, float h, const float* __restrict__ alpha,
                                                        const float* __restrict__ D, const float* __restrict__ lk,
                                                        const float* __restrict__ lsrc, const float* __restrict__ lbet,
                                                        float* lalpha, float* ld) {
    const uint32_t i = pid();
    if (i >= n) return;
    const float beta = alpha[i] / h, Di = D[i];
    const float lD = -lsrc[i] + beta * lk[i];
    const float lb = lbet[i] + Di * lk[i];
    ld[i] = Di > 0.0f ? lD : 0.0f;
    lalpha[i] += lb / h;
}

// la = s * lv (s = h or H), lk = 0 is set by the caller; generic helpers
__global__ void k_adj_scale4(uint32_t n, const float4* __restrict__ in, float s, float4* out) {
    const uint32_t i = pid();
    if (i < n) out[i] = st4(s * ld3(in[i]));
}
__global__ void k_adj_axpy4(uint32_t n, float s, const float4* __restrict__ in, float4* out) {  // out += s in
    const uint32_t i = pid();
    if (i < n) out[i] = st4(ld3(out[i]) + s * ld3(in[i]));
}
__global__ void k_adj_fill(float* p, float v, uint32_t n) {
    const uint32_t i = pid();
    if (i < n) p[i] = v;
}

// viscosity system T_ij = k_ij (r r^T)(y_i - y_j), k_ij = (h / rho_i) 10 mu m / (rho_j (|r|^2 + 0.01 R^2)) G(|r|)
__device__ __forceinline__ float visc_k(const AdjK& a, float h, float rhoi, float rhoj, float rl, float g) {
    return (h / rhoi) * (10.0f * a.mu * a.m / rhoj) * g / (rl * rl + 0.01f * a.R2);
}
// out = A^T w: out_i = w_i - sum_j r r^T (k_ij w_i - k_ji w_j) = w_i - sum_j k_ij r r^T (w_i - w_j), since
// k_ij = k_ji = h 10 mu m G / (rho_i rho_j (|r|^2 + 0.01 R^2))
struct AdjViscT {
    f3 xi, wi, o;
    float hci;  // h 10 mu m / rho_i
    __device__ __forceinline__ AdjViscT(const AdjK& a, uint32_t i, float h, const float* __restrict__ rho,
                                        const float4* __restrict__ w)
        : xi(ld3(a.x[i])), wi(ld3(w[i])), o(wi), hci(h * (10.0f * a.mu * a.m) * rcp(rho[i])) {}
    __device__ __forceinline__ void term(const AdjK& a, f3 xj, f3 wj, float rhoj, float g) {
        const f3 r = xi - xj;
        const float k = hci * rcp(rhoj) * g * rcp(dot3(r, r) + 0.01f * a.R2);
        o = o - (k * dot3(r, wi - wj)) * r;
    }
    __device__ __forceinline__ void finish(const AdjK& a, uint32_t i, float h, const float* __restrict__ rho,
                                           float4* out) {
        if (a.bodies && a.muB != 0.0f) o = o - (h / rho[i]) * adj_visc_boundary(a, i, wi);  // -(h/rho_i) C_i w_i
        out[i] = st4(o);
    }
};
__global__ void __launch_bounds__(kTile) k_adj_visc_T(AdjK a, float h, const float* __restrict__ rho,
                                                      const float4* __restrict__ w, float4* out,
                                                      const BicgState* __restrict__ st) {
    if (st && st->done) return;  // a device-side solve that has converged (block-uniform)
    uint32_t i;
    if (!adj_slot(a, i)) return;
    AdjViscT t(a, i, h, rho, w);
    const uint32_t cnt = a.count[i];
    for (uint32_t j0 = 0; j0 < cnt; j0 += kAdjBatch) {
        f3 xj[kAdjBatch], wj[kAdjBatch];
        float rj[kAdjBatch];
#pragma unroll
        for (uint32_t q = 0; q < kAdjBatch; ++q) {
            const uint32_t j = adj_nbr(a, i, min(j0 + q, cnt - 1u));
            xj[q] = ld3(a.x[j]);
            wj[q] = ld3(w[j]);
            rj[q] = rho[j];
        }
#pragma unroll
        for (uint32_t q = 0; q < kAdjBatch; ++q) {
            if (j0 + q >= cnt) break;
            const f3 r = t.xi - xj[q];
            t.term(a, xj[q], wj[q], rj[q], adj_g(a, sqrtf(dot3(r, r))));
        }
    }
    t.finish(a, i, h, rho, out);
}
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_adj_visc_T_groups(AdjK a, AdjStage g, float h,
                                                                                const float* __restrict__ rho,
                                                                                const float4* __restrict__ w,
                                                                                float4* out,
                                                                                const BicgState* __restrict__ st) {
    if (st && st->done) return;  // a device-side solve that has converged (block-uniform)
    adj_group_loop(
        a, g, rho, w, [&](uint32_t i) { return AdjViscT(a, i, h, rho, w); },
        [&](AdjViscT& t, uint32_t i) { t.finish(a, i, h, rho, out); });
}
// parameter cotangents of -lb^T (dA) y = lb . dT: positions, densities, and per-tile partial sums of lb . T / mu
__global__ void __launch_bounds__(kTile) k_adj_visc_vjp(AdjK a, float h, const float* __restrict__ rho,
                                                        const float4* __restrict__ y, const float4* __restrict__ lb,
                                                        float4* lx, float* lrho, double* muPart) {
    uint32_t i;
    double part = 0.0;
    if (adj_slot(a, i)) {
        const f3 xi = ld3(a.x[i]), yi = ld3(y[i]), lbi = ld3(lb[i]);
        const float e = 0.01f * a.R2;
        f3 gx = mk3(0.0f, 0.0f, 0.0f);
        float lr = 0.0f;
        for (uint32_t jj = 0; jj < a.count[i]; ++jj) {
            const uint32_t j = adj_nbr(a, i, jj);
            const f3 r = xi - ld3(a.x[j]);
            const float rl = sqrtf(dot3(r, r)), g = adj_g(a, rl);
            const f3 u = yi - ld3(y[j]), lbj = ld3(lb[j]);
            const float ur = dot3(u, r), si = dot3(lbi, r), sj = dot3(lbj, r);
            const float kij = visc_k(a, h, rho[i], rho[j], rl, g), kji = visc_k(a, h, rho[j], rho[i], rl, g);
            const float lTij = kij * ur * si, lTji = -kji * ur * sj;  // lb_i . T_ij, lb_j . T_ji
            part += (double)lTij;
            lr -= (lTij + lTji) / rho[i];
            const float d2 = rl * rl + e, dg = adj_dG(a, rl);
            const float Ci = (h / rho[i]) * (10.0f * a.mu * a.m / rho[j]);
            const float Cj = (h / rho[j]) * (10.0f * a.mu * a.m / rho[i]);
            const float dki = Ci * (dg * d2 - g * 2.0f * rl) / (d2 * d2);
            const float dkj = Cj * (dg * d2 - g * 2.0f * rl) / (d2 * d2);
            f3 t = kij * (ur * lbi + si * u) - kji * (ur * lbj + sj * u);
            if (rl > 0.0f) t = t + (ur * (si * dki - sj * dkj) / rl) * r;
            gx = gx + t;
        }
        if (a.bodies && a.muB != 0.0f) {  // boundary samples: lb_i . (h / rho_i) C_i(x_i) y_i
            const float hr = h / rho[i];
            lr -= hr * dot3(lbi, adj_visc_boundary(a, i, yi)) / rho[i];
            gx = gx + hr * adj_visc_boundary_vjp(a, i, lbi, yi);
        }
        lx[i] = st4(ld3(lx[i]) + gx);
        lrho[i] += lr;
    }
    double v[1] = {part};
    block_sum<1>(v);
    if (threadIdx.x == 0) muPart[blockIdx.x] = v[0] / (double)a.mu;
}

// factor: S_i = V^2 sum_j |gradW_ij|^2 + |gs_i|^2, gs_i = V sum_j gradW_ij + B_i, alpha = 1/S
__global__ void __launch_bounds__(kTile) k_adj_gsum(AdjK a, float4* gs) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    const f3 xi = ld3(a.x[i]);
    f3 s = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t jj = 0; jj < a.count[i]; ++jj) {
        const f3 r = xi - ld3(a.x[adj_nbr(a, i, jj)]);
        s = s + adj_g(a, sqrtf(dot3(r, r))) * r;
    }
    s = a.V * s;
    if (a.bodies) s = s + adj_contact_sum(a, i);
    gs[i] = st4(s);
}
__global__ void __launch_bounds__(kTile) k_adj_factor_lS(uint32_t n, const float* __restrict__ alpha,
                                                         const float* __restrict__ lalpha, float* lS) {
    const uint32_t i = pid();
    if (i < n) lS[i] = alpha[i] == 0.0f ? 0.0f : -alpha[i] * alpha[i] * lalpha[i];
}
// lx_i += sum_j vjp(r_ij, 2V [V gradW_ij (lS_i + lS_j) + lS_i gs_i - lS_j gs_j])
//       + sum_j rho0 V (lrho_i + lrho_j) W'(|r|) r / |r|                     (densities)
//       + d/dx_i (2 lS_i gs_i . B_i + rho0 lrho_i sum_b vb W(x_i - xb))       (boundary samples)
__global__ void __launch_bounds__(kTile) k_adj_factor_density(AdjK a, const float* __restrict__ lS,
                                                              const float4* __restrict__ gs,
                                                              const float* __restrict__ lrho, float4* lx) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    const f3 xi = ld3(a.x[i]), gsi = ld3(gs[i]);
    const float lSi = lS[i], lri = lrho[i];
    f3 gx = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t jj = 0; jj < a.count[i]; ++jj) {
        const uint32_t j = adj_nbr(a, i, jj);
        const f3 r = xi - ld3(a.x[j]);
        const float rl = sqrtf(dot3(r, r)), g = adj_g(a, rl);
        const f3 c = (2.0f * a.V) * ((a.V * (lSi + lS[j])) * (g * r) + lSi * gsi - lS[j] * ld3(gs[j]));
        gx = gx + vjp_grad(a, r, rl, g, c);
        if (rl > 0.0f) gx = gx + (a.rho0 * a.V * (lri + lrho[j]) * adj_dW(a, rl) / rl) * r;
    }
    if (a.bodies)
        gx = gx + adj_contact_vjp(a, i, (2.0f * lSi) * gsi) + (a.rho0 * lri) * adj_contact_density_grad(a, i);
    lx[i] = st4(ld3(lx[i]) + gx);
}

// Surface tension of the step with the Monte-Carlo normals nMC and curvatures kMC held (piecewise constant in the
// positions): for nMC_i != 0, w_ij = 1 - |x_i - x_j| / R over neighbours with nMC_j != 0,
//   nc = sum w nMC_j, cc = sum w kMC_j, cf = sum w, nS = normalize((1 - sf) nMC_i + sf normalize(nc)),
//   kS = ((1 - sf) kMC_i + sf cc) / (1 - sf + sf cf), a_i = c kS nS with c = -(passes sigma / m).
// Pointwise part: from la = s lv, the cotangents of nc, cc and cf (zero where nMC_i = 0).
__device__ __forceinline__ f3 adj_normalize_T(f3 a, f3 l) {  // d normalize(a)^T l; identity where a = 0
    const float a2 = dot3(a, a);
    if (!(a2 > 0.0f)) return l;
    const f3 u = normalize(a);
    return (l - dot3(u, l) * u) / sqrtf(a2);
}
// The force is -(sigma / m) E nS with E the sum of the P passes' curvatures: E = P kS, or with temporal smoothing
// (dK0 != null: the curvature each particle carried in) E = A dK0 + (P - A) kS, A = sum_(p=1..P) 0.95^p, and the
// particle carries out dK' = aP dK0 + (1 - aP) kS, aP = 0.95^P (0.95^P dK0 off the surface); without it dK' = kS (0 off
// the surface). ldK (null: not tracked) holds the cotangent of dK' on entry and of dK on return.
__global__ void __launch_bounds__(kTile) k_adj_st_point(AdjK a, float s, const float4* __restrict__ lv,
                                                        const float4* __restrict__ nMC, const float* __restrict__ kMC,
                                                        float sf, float c, float P, float A, float aP,
                                                        const float* __restrict__ dK0, float* ldK, float4* lnc,
                                                        float2* lcw) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    const f3 ni = ld3(nMC[i]);
    if (is_zero(ni)) {
        lnc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        lcw[i] = make_float2(0.0f, 0.0f);
        if (ldK) ldK[i] = dK0 ? aP * ldK[i] : 0.0f;
        return;
    }
    const f3 xi = ld3(a.x[i]);
    f3 nc = mk3(0.0f, 0.0f, 0.0f);
    float cc = 0.0f, cf = 0.0f;
    for (uint32_t jj = 0; jj < a.count[i]; ++jj) {
        const uint32_t j = adj_nbr(a, i, jj);
        const f3 nj = ld3(nMC[j]);
        if (is_zero(nj)) continue;
        const float w = 1.0f - length(ld3(a.x[j]) - xi) / a.R;
        nc = nc + nj * w;
        cc += kMC[j] * w;
        cf += w;
    }
    const f3 m = dot3(nc, nc) > 0.0f ? normalize(nc) : nc;
    const f3 u = (1.0f - sf) * ni + sf * m;
    const f3 ns = dot3(u, u) > 0.0f ? normalize(u) : u;
    const float den = 1.0f - sf + sf * cf;
    const float ks = ((1.0f - sf) * kMC[i] + sf * cc) / den;
    const f3 la = s * ld3(lv[i]);
    const float lan = dot3(la, ns), l = ldK ? ldK[i] : 0.0f;
    float E, lks;
    if (dK0) {
        E = A * dK0[i] + (P - A) * ks;
        lks = c * (P - A) * lan + (1.0f - aP) * l;
        if (ldK) ldK[i] = c * A * lan + aP * l;
    } else {
        E = P * ks;
        lks = c * P * lan + l;
        if (ldK) ldK[i] = 0.0f;
    }
    lnc[i] = st4(adj_normalize_T(nc, sf * adj_normalize_T(u, (c * E) * la)));
    lcw[i] = make_float2(lks * sf / den, -lks * ks * sf / den);  // cotangents of cc and cf
}
// Pair part: lx_i -= sum_j (lw_ij + lw_ji) (x_i - x_j) / (R |x_i - x_j|) over pairs with both nMC != 0,
// lw_ij = lnc_i . nMC_j + lcc_i kMC_j + lcf_i
__global__ void __launch_bounds__(kTile) k_adj_st_pairs(AdjK a, const float4* __restrict__ nMC,
                                                        const float* __restrict__ kMC, const float4* __restrict__ lnc,
                                                        const float2* __restrict__ lcw, float4* lx) {
    uint32_t i;
    if (!adj_slot(a, i)) return;
    const f3 ni = ld3(nMC[i]);
    if (is_zero(ni)) return;
    const f3 xi = ld3(a.x[i]), li = ld3(lnc[i]);
    const float ki = kMC[i];
    const float2 ci = lcw[i];
    f3 gx = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t jj = 0; jj < a.count[i]; ++jj) {
        const uint32_t j = adj_nbr(a, i, jj);
        const f3 nj = ld3(nMC[j]);
        if (is_zero(nj)) continue;
        const f3 d = xi - ld3(a.x[j]);
        const float dl = length(d);
        if (!(dl > 0.0f)) continue;
        const float2 cj = lcw[j];
        const float lw = (dot3(li, nj) + ci.x * kMC[j] + ci.y) + (dot3(ld3(lnc[j]), ni) + cj.x * ki + cj.y);
        gx = gx - (lw / (a.R * dl)) * d;
    }
    lx[i] = st4(ld3(lx[i]) + gx);
}

// push-out reversed: x = x_before + delta n(x_before), v = 0  ->  lx = J^T lx, lv = 0 (rows of J in
// push[q stride + i])
__global__ void __launch_bounds__(kTile) k_adj_push(uint32_t n, uint32_t stride, const float4* __restrict__ push,
                                                    float4* lx, float4* lv) {
    const uint32_t i = pid();
    if (i >= n) return;
    const float4 r0 = push[i];
    if (!(r0.w > 0.0f)) return;
    const f3 l = ld3(lx[i]);
    lx[i] = st4(l.x * mk3(r0.x, r0.y, r0.z) + l.y * ld3(push[stride + i]) + l.z * ld3(push[2 * (size_t)stride + i]));
    lv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---- BiCGSTAB with its scalars on the device --------------------------------------------------------------------
// The solve of the transposed viscosity system runs a fixed sequence of launches per iteration; the scalars live in
// one BicgState, every kernel returns at once when `done` is set, and the host reads `done` only every few
// iterations. Dots are FP64 per-block partials summed in a fixed order, scalar updates are in double and the vector
// updates round the scalars to float.
__global__ void __launch_bounds__(kTile) k_bicg_dot(uint32_t n, const float4* __restrict__ p,
                                                    const float4* __restrict__ q, const float4* __restrict__ p2,
                                                    const float4* __restrict__ q2, double* part,
                                                    const BicgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = pid();
    double v[2] = {i < n ? (double)dot3(ld3(p[i]), ld3(q[i])) : 0.0,
                   (p2 && i < n) ? (double)dot3(ld3(p2[i]), ld3(q2[i])) : 0.0};
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v[0];
        part[gridDim.x + blockIdx.x] = v[1];
    }
}
// the dots' block partials summed by the block in a fixed order (strided per thread, then block_sum's tree), then
// the scalar update of stage `which`
__global__ void __launch_bounds__(kTile) k_bicg_scalars(const double* __restrict__ part, uint32_t blocks,
                                                        BicgState* st, int which) {
    if (st->done) return;  // block-uniform
    double v[2] = {0.0, 0.0};
    for (uint32_t k = threadIdx.x; k < blocks; k += blockDim.x) {
        v[0] += part[k];
        if (which == kBicgTsTt || which == kBicgBegin) v[1] += part[blocks + k];
    }
    block_sum<2>(v);
    if (threadIdx.x != 0) return;
    const double a = v[0], b = v[1];
    switch (which) {
        case kBicgBegin:  // a = |b|^2, b = |r0|^2
            st->rr = b;
            st->tol2 = 1e-12 * a;
            if (a == 0.0 || !(b > st->tol2)) st->done = 1u;
            break;
        case kBicgRho:
            st->rhoNew = a;
            st->beta = (a / st->rhoOld) * (st->alpha / st->omega);
            break;
        case kBicgRv:
            st->alpha = st->rhoNew / a;
            break;
        case kBicgTsTt:
            st->omega = b > 0.0 ? a / b : 0.0;
            break;
        default: {  // kBicgRr: end of an iteration
            st->rr = a;
            st->rhoOld = st->rhoNew;
            st->it += 1;
            if (st->omega == 0.0 || !(st->it < st->maxIt && a > st->tol2)) st->done = 1u;
        }
    }
}
__global__ void k_bicg_reset(BicgState* st, uint32_t maxIt) {
    if (threadIdx.x == 0) *st = BicgState{1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0u, maxIt, 0u, 0u};
}
// acc[0] += sum of part[0, blocks): strided per thread, then block_sum's tree
__global__ void __launch_bounds__(kTile) k_adj_mu_acc(const double* __restrict__ part, uint32_t blocks, double* acc) {
    double v[1] = {0.0};
    for (uint32_t k = threadIdx.x; k < blocks; k += blockDim.x) v[0] += part[k];
    block_sum<1>(v);
    if (threadIdx.x == 0) acc[0] += v[0];
}
// p = r + beta (p - omega v)
__global__ void k_bicg_p(uint32_t n, const float4* __restrict__ r, const float4* __restrict__ v, float4* p,
                         const BicgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = pid();
    const float beta = (float)st->beta, omega = (float)st->omega;
    if (i < n) p[i] = st4(ld3(r[i]) + beta * (ld3(p[i]) - omega * ld3(v[i])));
}
// s = r - alpha v
__global__ void k_bicg_s(uint32_t n, const float4* __restrict__ r, const float4* __restrict__ v, float4* s,
                         const BicgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = pid();
    const float na = (float)-st->alpha;
    if (i < n) s[i] = st4(ld3(r[i]) + na * ld3(v[i]));
}
// x += alpha p, then x += omega s; r = s - omega t
__global__ void k_bicg_xr(uint32_t n, const float4* __restrict__ p, const float4* __restrict__ s,
                          const float4* __restrict__ t, float4* x, float4* r, const BicgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = pid();
    const float alpha = (float)st->alpha, omega = (float)st->omega, no = (float)-st->omega;
    if (i < n) {
        const f3 si = ld3(s[i]);
        x[i] = st4((ld3(x[i]) + alpha * ld3(p[i])) + omega * si);
        r[i] = st4(si + no * ld3(t[i]));
    }
}

// three copies in one launch (a recorded Jacobi iteration)
__global__ void k_adj_copy3(uint32_t n, const float* __restrict__ a, float* da, const float4* __restrict__ b,
                            float4* db, const float* __restrict__ c, float* dc) {
    const uint32_t i = pid();
    if (i < n) {
        da[i] = a[i];
        db[i] = b[i];
        dc[i] = c[i];
    }
}

// cotangents between the step's slot order and the original particle order (three floats per particle, or one)
__global__ void k_adj_to_slots1(uint32_t n, const uint32_t* __restrict__ id, const float* __restrict__ orig,
                                float* slot) {
    const uint32_t i = pid();
    if (i < n) slot[i] = orig[id[i]];
}
__global__ void k_adj_to_orig1(uint32_t n, const uint32_t* __restrict__ id, const float* __restrict__ slot,
                               float* orig) {
    const uint32_t i = pid();
    if (i < n) orig[id[i]] = slot[i];
}
__global__ void k_adj_to_slots(uint32_t n, const uint32_t* __restrict__ id, const float* __restrict__ orig,
                               float4* slot) {
    const uint32_t i = pid();
    if (i < n) {
        const uint32_t o = id[i];
        slot[i] = make_float4(orig[3 * o], orig[3 * o + 1], orig[3 * o + 2], 0.0f);
    }
}
__global__ void k_adj_to_orig(uint32_t n, const uint32_t* __restrict__ id, const float4* __restrict__ slot,
                              float* orig) {
    const uint32_t i = pid();
    if (i < n) {
        const uint32_t o = id[i];
        const float4 v = slot[i];
        orig[3 * o] = v.x;
        orig[3 * o + 1] = v.y;
        orig[3 * o + 2] = v.z;
    }
}

// ---- host side ---------------------------------------------------------------------------------------------------
static inline dim3 grid_of(uint32_t n) { return dim3((n + kTile - 1) / kTile); }

static inline AdjStage adj_stage(const AdjKHost& h) {
    AdjStage g;
    g.tg.tiles = h.groups.tiles;
    g.tg.per = h.groups.per;
    g.H = HaloView{h.groups.haloCount, h.groups.haloIds};
    g.local = h.groups.local;
    return g;
}
void adj_launch_accel_T(hipStream_t s, const AdjKHost& h, const float* kap, const float4* la, float* lk, float4* lx) {
    if (h.groups.grid)
        hipLaunchKernelGGL(k_adj_accel_T_groups, dim3(h.groups.grid), dim3(kGroupSize), kLdsAdjStage, s, h.dev(),
                           adj_stage(h), kap, la, lk, lx);
    else
        hipLaunchKernelGGL(k_adj_accel_T, grid_of(h.n), dim3(kTile), 0, s, h.dev(), kap, la, lk, lx);
}
void adj_launch_div_T(hipStream_t s, const AdjKHost& h, const float* lf, const float4* u, float4* lu, float4* lx) {
    if (h.groups.grid)
        hipLaunchKernelGGL(k_adj_div_T_groups, dim3(h.groups.grid), dim3(kGroupSize), kLdsAdjStage, s, h.dev(),
                           adj_stage(h), lf, u, lu, lx);
    else
        hipLaunchKernelGGL(k_adj_div_T, grid_of(h.n), dim3(kTile), 0, s, h.dev(), lf, u, lu, lx);
}
int configure_adjoint_kernels() {
    for (const void* fn :
         {(const void*)k_adj_accel_T_groups, (const void*)k_adj_div_T_groups, (const void*)k_adj_visc_T_groups})
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsAdjStage) != hipSuccess)
            return 1;
    return 0;
}
void adj_launch_jacobi_point(hipStream_t s, uint32_t n, const float* lk, const float* alpha, float div,
                             const float* res, float c, float* lsrc, float* ls, float* lf) {
    hipLaunchKernelGGL(k_adj_jacobi_point, grid_of(n), dim3(kTile), 0, s, n, lk, alpha, div, res, c, lsrc, ls, lf);
}
void adj_launch_pres_init(hipStream_t s, uint32_t n, float H, float rho0, const float* alpha, const float* rhoadv,
                          const float* lk, const float* lsrc, const float* lgam, float* lalpha, float* lrho,
                          float* ld2) {
    hipLaunchKernelGGL(k_adj_pres_init, grid_of(n), dim3(kTile), 0, s, n, H, rho0, alpha, rhoadv, lk, lsrc, lgam,
                       lalpha, lrho, ld2);
}
void adj_launch_div_init(hipStream_t s, uint32_t n, float h, const float* alpha, const float* D, const float* lk,
                         const float* lsrc, const float* lbet, float* lalpha, float* ld) {
    hipLaunchKernelGGL(k_adj_div_init, grid_of(n), dim3(kTile), 0, s, n, h, alpha, D, lk, lsrc, lbet, lalpha, ld);
}
void adj_launch_scale4(hipStream_t s, uint32_t n, const float4* in, float f, float4* out) {
    hipLaunchKernelGGL(k_adj_scale4, grid_of(n), dim3(kTile), 0, s, n, in, f, out);
}
void adj_launch_axpy4(hipStream_t s, uint32_t n, float f, const float4* in, float4* out) {
    hipLaunchKernelGGL(k_adj_axpy4, grid_of(n), dim3(kTile), 0, s, n, f, in, out);
}
void adj_launch_fill(hipStream_t s, float* p, float v, uint32_t n) {
    hipLaunchKernelGGL(k_adj_fill, grid_of(n), dim3(kTile), 0, s, p, v, n);
}
void adj_launch_visc_T(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const float4* w, float4* out,
                       const BicgState* st) {
    if (h.groups.grid)
        hipLaunchKernelGGL(k_adj_visc_T_groups, dim3(h.groups.grid), dim3(kGroupSize), kLdsAdjStage, s, h.dev(),
                           adj_stage(h), dt, rho, w, out, st);
    else
        hipLaunchKernelGGL(k_adj_visc_T, grid_of(h.n), dim3(kTile), 0, s, h.dev(), dt, rho, w, out, st);
}
void adj_launch_visc_vjp(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const float4* y,
                         const float4* lb, float4* lx, float* lrho, double* muPart) {
    hipLaunchKernelGGL(k_adj_visc_vjp, grid_of(h.n), dim3(kTile), 0, s, h.dev(), dt, rho, y, lb, lx, lrho, muPart);
}
void adj_launch_factor_density(hipStream_t s, const AdjKHost& h, const float* alpha, const float* lalpha, float* lS,
                               float4* gs, const float* lrho, float4* lx) {
    hipLaunchKernelGGL(k_adj_gsum, grid_of(h.n), dim3(kTile), 0, s, h.dev(), gs);
    hipLaunchKernelGGL(k_adj_factor_lS, grid_of(h.n), dim3(kTile), 0, s, h.n, alpha + h.base, lalpha + h.base,
                       lS + h.base);
    hipLaunchKernelGGL(k_adj_factor_density, grid_of(h.n), dim3(kTile), 0, s, h.dev(), lS, gs, lrho, lx);
}
void adj_launch_copy3(hipStream_t s, uint32_t n, const float* a, float* da, const float4* b, float4* db,
                      const float* c, float* dc) {
    if (n) hipLaunchKernelGGL(k_adj_copy3, grid_of(n), dim3(kTile), 0, s, n, a, da, b, db, c, dc);
}
void adj_launch_to_slots(hipStream_t s, uint32_t n, const uint32_t* id, const float* orig, float4* slot) {
    hipLaunchKernelGGL(k_adj_to_slots, grid_of(n), dim3(kTile), 0, s, n, id, orig, slot);
}
void adj_launch_to_orig(hipStream_t s, uint32_t n, const uint32_t* id, const float4* slot, float* orig) {
    hipLaunchKernelGGL(k_adj_to_orig, grid_of(n), dim3(kTile), 0, s, n, id, slot, orig);
}
void adj_launch_to_slots1(hipStream_t s, uint32_t n, const uint32_t* id, const float* orig, float* slot) {
    hipLaunchKernelGGL(k_adj_to_slots1, grid_of(n), dim3(kTile), 0, s, n, id, orig, slot);
}
void adj_launch_to_orig1(hipStream_t s, uint32_t n, const uint32_t* id, const float* slot, float* orig) {
    hipLaunchKernelGGL(k_adj_to_orig1, grid_of(n), dim3(kTile), 0, s, n, id, slot, orig);
}

AdjK AdjKHost::dev() const {
    return AdjK{tab,   R,    invStep, K,     L,    V,     rho0, m,   mu,   R2,     n,   cap,
                base,  stride ? stride : n,   count, list, x,     con,  bodies, muB, td, 1.0f / R,
                3.0f * L / (R * R * R), L / R};
}

void adj_launch_surface_tension(hipStream_t s, const AdjKHost& h, float scale, const float4* lv, const float4* nMC,
                                const float* kMC, float sf, float c, uint32_t passes, const float* dK0, float* ldK,
                                float4* lnc, float2* lcw, float4* lx) {
    float A = 0.0f, aP = 1.0f;
    for (uint32_t q = 0; q < passes; ++q) {
        aP *= 0.95f;
        A += aP;
    }
    hipLaunchKernelGGL(k_adj_st_point, grid_of(h.n), dim3(kTile), 0, s, h.dev(), scale, lv, nMC, kMC, sf, c,
                       (float)passes, A, aP, dK0, ldK, lnc, lcw);
    hipLaunchKernelGGL(k_adj_st_pairs, grid_of(h.n), dim3(kTile), 0, s, h.dev(), nMC, kMC, lnc, lcw, lx);
}
void adj_launch_push(hipStream_t s, uint32_t n, uint32_t stride, const float4* push, float4* lx, float4* lv) {
    if (n) hipLaunchKernelGGL(k_adj_push, grid_of(n), dim3(kTile), 0, s, n, stride, push, lx, lv);
}

void adj_launch_bicg_begin(hipStream_t s, const AdjKHost& h, const float4* b, const float4* r, double* part,
                           BicgState* st, uint32_t maxIt) {
    const uint32_t n = h.n, o = h.base;
    hipLaunchKernelGGL(k_bicg_reset, dim3(1), dim3(kWave), 0, s, st, maxIt);
    hipLaunchKernelGGL(k_bicg_dot, grid_of(n), dim3(kTile), 0, s, n, b + o, b + o, r + o, r + o, part, st);
    hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(kTile), 0, s, part, grid_of(n).x, st, (int)kBicgBegin);
}
void adj_launch_mu_acc(hipStream_t s, const double* part, uint32_t blocks, double* acc) {
    hipLaunchKernelGGL(k_adj_mu_acc, dim3(1), dim3(kTile), 0, s, part, blocks, acc);
}
void adj_launch_bicg_iteration(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const BicgVectors& v,
                               double* part, BicgState* st) {
    // the vectors are slot-indexed arrays: the operator takes them whole, the pointwise kernels from the scene's
    // first slot
    const uint32_t n = h.n, b = h.base;
    const uint32_t blocks = grid_of(n).x;
    hipLaunchKernelGGL(k_bicg_dot, grid_of(n), dim3(kTile), 0, s, n, v.rh + b, v.r + b, nullptr, nullptr, part, st);
    hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(kTile), 0, s, part, blocks, st, (int)kBicgRho);
    hipLaunchKernelGGL(k_bicg_p, grid_of(n), dim3(kTile), 0, s, n, v.r + b, v.v + b, v.p + b, st);
    adj_launch_visc_T(s, h, dt, rho, v.p, v.v, st);
    hipLaunchKernelGGL(k_bicg_dot, grid_of(n), dim3(kTile), 0, s, n, v.rh + b, v.v + b, nullptr, nullptr, part, st);
    hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(kTile), 0, s, part, blocks, st, (int)kBicgRv);
    hipLaunchKernelGGL(k_bicg_s, grid_of(n), dim3(kTile), 0, s, n, v.r + b, v.v + b, v.s + b, st);
    adj_launch_visc_T(s, h, dt, rho, v.s, v.t, st);
    hipLaunchKernelGGL(k_bicg_dot, grid_of(n), dim3(kTile), 0, s, n, v.t + b, v.s + b, v.t + b, v.t + b, part, st);
    hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(kTile), 0, s, part, blocks, st, (int)kBicgTsTt);
    hipLaunchKernelGGL(k_bicg_xr, grid_of(n), dim3(kTile), 0, s, n, v.p + b, v.s + b, v.t + b, v.x + b, v.r + b, st);
    hipLaunchKernelGGL(k_bicg_dot, grid_of(n), dim3(kTile), 0, s, n, v.r + b, v.r + b, nullptr, nullptr, part, st);
    hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(kTile), 0, s, part, blocks, st, (int)kBicgRr);
}

}  // namespace vfd
