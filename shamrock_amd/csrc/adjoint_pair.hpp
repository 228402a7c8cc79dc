// SPDX-License-Identifier: CeCILL-2.1
//
// adjoint_pair.hpp -- reverse-mode derivatives of the SPH pair terms of one step, written by
// hand (no reference counterpart: the reference has no adjoint).  Host and device code, so the
// pair-level derivatives are checked against finite differences on the CPU.
//
//   force pair (UpdateDerivs.cpp:88-288, constant AV, adiabatic or isothermal EOS): the
//     contribution of neighbour b to a_a and du_a as a function of r = x_a - x_b, v_a, v_b,
//     u_a, u_b, h_a, h_b, Omega_a, Omega_b (rho = rho_h(h), P and c_s from the EOS);
//   density / Omega pair terms: m W(r, h_a), m dW/dh(r, h_a) and their derivatives.
//
// Kinks are differentiated a.e.: max(q, 0) and |v.r| take the branch of the forward value, the
// conductivity speed sqrt(|dP| / rho_avg) has derivative 0 where dP = 0, and a pair closer than
// 1e-9 (where the forward sets the unit vector to 0 and every term vanishes) has none.
#pragma once

#include "common.hpp"

#include <cmath>

namespace shamx {

    // second derivative f'' of the kernel profile f(q) (sphkernels.hpp polynomials)
    template<class K>
    struct KernD2;
    template<>
    struct KernD2<M4> {
        __host__ __device__ static f64 ddf(f64 q) {
            if (q < 1) return -3 + 4.5 * q;
            if (q < 2) return 3 - 1.5 * q;
            return 0;
        }
    };
    template<>
    struct KernD2<M6> {
        __host__ __device__ static f64 ddf(f64 q) {
            f64 t1 = 3 - q, t2 = 2 - q, t3 = 1 - q;
            f64 a = 20 * t1 * t1 * t1, b = -120 * t2 * t2 * t2, c = 300 * t3 * t3 * t3;
            if (q < 1) return a + b + c;
            if (q < 2) return a + b;
            if (q < 3) return a;
            return 0;
        }
    };
    template<>
    struct KernD2<M8> {
        __host__ __device__ static f64 ddf(f64 q) {
            auto p5 = [](f64 v) { return v * v * v * v * v; };
            f64 a = 42 * p5(4 - q), b = -336 * p5(3 - q), c = 1176 * p5(2 - q),
                d = -2352 * p5(1 - q);
            if (q < 1) return a + b + c + d;
            if (q < 2) return a + b + c;
            if (q < 3) return a + b;
            if (q < 4) return a;
            return 0;
        }
    };
    template<>
    struct KernD2<C2> {
        __host__ __device__ static f64 ddf(f64 q) {
            f64 p1 = 1 - 0.5 * q;
            if (q < 2) return p1 * p1 * (10 * q - 5);
            return 0;
        }
    };
    template<class K>
    constexpr bool has_adjoint() {
        return K::id == M4::id || K::id == M6::id || K::id == M8::id || K::id == C2::id;
    }

    // W(r, h) and its first and second partial derivatives in r and h
    template<class K>
    struct KernDeriv {
        static constexpr f64 norm = K::norm_3d;
        f64 W, Wr, Wh, Wrr, Wrh, Whh;
        __host__ __device__ KernDeriv(f64 r, f64 h) {
            f64 q = r / h, f = K::f(q), df = K::df(q), ddf = KernD2<K>::ddf(q);
            f64 h2 = h * h, h3 = h2 * h, h4 = h3 * h, h5 = h4 * h;
            W   = norm * f / h3;
            Wr  = norm * df / h4;
            Wh  = -norm * (3 * f + q * df) / h4;
            Wrr = norm * ddf / h5;
            Wrh = -norm * (4 * df + q * ddf) / h5;
            Whh = norm * (12 * f + 8 * q * df + q * q * ddf) / h5;
        }
    };

    struct PairParams {
        f64 mass, hfact, gamma, alpha_u, alpha_AV, beta_AV;
        int eos_kind; // 0 adiabatic, 1 isothermal (sound speed cs_iso), 2 LP07 (side's cs)
        f64 cs_iso;
    };

    // per-particle state entering a force pair (cs: the particle's LP07 sound speed, which
    // depends on its position; unused by the other equations of state)
    struct PairSide {
        f64 v[3], u, h, omega, cs;
    };
    // gradient of the seeded pair outputs w.r.t. one side's inputs
    struct PairGrad {
        f64 v[3], u, h, omega, cs;
    };

    // P, c_s and their partial derivatives; for LP07 (ComputeEos.cpp:726-798) P = c_s^2 rho
    // with c_s the side's own (dP_dcs = 2 c_s rho, c_s passes through)
    __host__ __device__ inline void eos_of(const PairParams &p, f64 rho, f64 u, f64 cs_loc,
                                           f64 &P, f64 &cs, f64 &dP_drho, f64 &dP_du,
                                           f64 &dcs_drho, f64 &dcs_du, f64 &dP_dcs) {
        dP_dcs = 0;
        if (p.eos_kind == 2) {
            P        = cs_loc * cs_loc * rho;
            cs       = cs_loc;
            dP_drho  = cs_loc * cs_loc;
            dP_du    = 0;
            dcs_drho = 0;
            dcs_du   = 0;
            dP_dcs   = 2 * cs_loc * rho;
        } else if (p.eos_kind == 0) {
            P  = (p.gamma - 1) * rho * u;
            cs = std::sqrt(p.gamma * P / rho);
            dP_drho = (p.gamma - 1) * u;
            dP_du   = (p.gamma - 1) * rho;
            // c_s = sqrt(gamma (gamma - 1) u): independent of rho
            dcs_drho = 0;
            dcs_du   = cs > 0 ? p.gamma * (p.gamma - 1) / (2 * cs) : 0;
        } else {
            P        = p.cs_iso * p.cs_iso * rho;
            cs       = p.cs_iso;
            dP_drho  = p.cs_iso * p.cs_iso;
            dP_du    = 0;
            dcs_drho = 0;
            dcs_du   = 0;
        }
    }

    // Contribution of neighbour b to (a_a, du_a) (forward), or, with seeds (ab[3], db) for
    // (a_a, du_a), the reverse sweep: gradients w.r.t. r = x_a - x_b (gr), side a and side b.
    template<class K>
    __host__ __device__ inline void force_pair(const PairParams &p, const f64 r[3],
                                               const PairSide &A, const PairSide &B, f64 acc[3],
                                               f64 &du, const f64 *ab, f64 db, f64 gr[3],
                                               PairGrad &gA, PairGrad &gB) {
        const f64 m = p.mass;
        // ---- forward
        f64 d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        f64 d  = std::sqrt(d2);
        f64 inv = (d >= 1e-9) ? 1. / d : 0.;
        f64 e[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
        f64 hfa = p.hfact / A.h, hfb = p.hfact / B.h;
        f64 rho_a = m * hfa * hfa * hfa, rho_b = m * hfb * hfb * hfb;
        f64 Pa, csa, dPa_r, dPa_u, dca_r, dca_u, dPa_c, Pb, csb, dPb_r, dPb_u, dcb_r, dcb_u, dPb_c;
        eos_of(p, rho_a, A.u, A.cs, Pa, csa, dPa_r, dPa_u, dca_r, dca_u, dPa_c);
        eos_of(p, rho_b, B.u, B.cs, Pb, csb, dPb_r, dPb_u, dcb_r, dcb_u, dPb_c);
        KernDeriv<K> ka(d, A.h), kb(d, B.h);
        f64 Fa = ka.Wr, Fb = kb.Wr;
        f64 vab[3] = {A.v[0] - B.v[0], A.v[1] - B.v[1], A.v[2] - B.v[2]};
        f64 w  = vab[0] * e[0] + vab[1] * e[1] + vab[2] * e[2];
        f64 aw = std::fabs(w);
        f64 vsa = p.alpha_AV * csa + p.beta_AV * aw;
        f64 vsb = p.alpha_AV * csb + p.beta_AV * aw;
        f64 ravg = (rho_a + rho_b) * 0.5;
        f64 dP   = Pa - Pb;
        f64 vsu  = std::sqrt(std::fabs(dP) / ravg);
        f64 qa_in = -0.5 * rho_a * vsa * w, qb_in = -0.5 * rho_b * vsb * w;
        f64 qa = qa_in > 0 ? qa_in : 0, qb = qb_in > 0 ? qb_in : 0;
        f64 APa = Pa + qa, APb = Pb + qb;
        f64 isa = 1. / (rho_a * rho_a * A.omega), isb = 1. / (rho_b * rho_b * B.omega);
        f64 Sa = APa * isa, Sb = APb * isb;
        f64 G  = Sa * Fa + Sb * Fb;
        f64 Ya = Fa / (rho_a * A.omega), Yb = Fb / (rho_b * B.omega);
        f64 Y  = Ya + Yb;
        f64 cu = m * p.alpha_u * 0.5;
        f64 dU = A.u - B.u;
        for (int k = 0; k < 3; k++) acc[k] = -m * G * e[k];
        du = m * Sa * Fa * w + cu * vsu * dU * Y;
        if (!ab) return;
        // ---- reverse
        gA = PairGrad{{0, 0, 0}, 0, 0, 0, 0};
        gB = PairGrad{{0, 0, 0}, 0, 0, 0, 0};
        gr[0] = gr[1] = gr[2] = 0;
        if (d < 1e-9) return;
        f64 ae    = ab[0] * e[0] + ab[1] * e[1] + ab[2] * e[2];
        f64 e_b[3] = {-m * G * ab[0], -m * G * ab[1], -m * G * ab[2]};
        f64 G_b   = -m * ae;
        // du = m Sa Fa w + cu vsu dU Y
        f64 Sa_b = db * m * Fa * w, Fa_b = db * m * Sa * w, w_b = db * m * Sa * Fa;
        f64 vsu_b = db * cu * dU * Y;
        f64 dU_b  = db * cu * vsu * Y;
        f64 Y_b   = db * cu * vsu * dU;
        gA.u += dU_b;
        gB.u -= dU_b;
        f64 rho_a_b = 0, rho_b_b = 0, Fb_b = 0, Sb_b = 0;
        Fa_b += Y_b / (rho_a * A.omega);
        rho_a_b += -Y_b * Ya / rho_a;
        gA.omega += -Y_b * Ya / A.omega;
        Fb_b += Y_b / (rho_b * B.omega);
        rho_b_b += -Y_b * Yb / rho_b;
        gB.omega += -Y_b * Yb / B.omega;
        // G = Sa Fa + Sb Fb
        Sa_b += G_b * Fa;
        Fa_b += G_b * Sa;
        Sb_b += G_b * Fb;
        Fb_b += G_b * Sb;
        // S = AP / (rho^2 Omega)
        f64 APa_b = Sa_b * isa, APb_b = Sb_b * isb;
        rho_a_b += -2 * Sa_b * Sa / rho_a;
        gA.omega += -Sa_b * Sa / A.omega;
        rho_b_b += -2 * Sb_b * Sb / rho_b;
        gB.omega += -Sb_b * Sb / B.omega;
        // AP = P + q
        f64 Pa_b = APa_b, Pb_b = APb_b, qa_b = APa_b, qb_b = APb_b;
        f64 vsa_b = 0, vsb_b = 0;
        if (qa_in > 0) {
            rho_a_b += qa_b * (-0.5 * vsa * w);
            vsa_b += qa_b * (-0.5 * rho_a * w);
            w_b += qa_b * (-0.5 * rho_a * vsa);
        }
        if (qb_in > 0) {
            rho_b_b += qb_b * (-0.5 * vsb * w);
            vsb_b += qb_b * (-0.5 * rho_b * w);
            w_b += qb_b * (-0.5 * rho_b * vsb);
        }
        // vsu = sqrt(|dP| / ravg)
        if (vsu > 0) {
            f64 z_b = vsu_b / (2 * vsu);
            f64 s   = dP > 0 ? 1. : -1.;
            f64 dP_b = z_b * s / ravg;
            Pa_b += dP_b;
            Pb_b -= dP_b;
            f64 ravg_b = -z_b * std::fabs(dP) / (ravg * ravg);
            rho_a_b += 0.5 * ravg_b;
            rho_b_b += 0.5 * ravg_b;
        }
        // vs = alpha cs + beta |w|
        f64 csa_b = p.alpha_AV * vsa_b, csb_b = p.alpha_AV * vsb_b;
        f64 aw_b  = p.beta_AV * (vsa_b + vsb_b);
        w_b += aw_b * (w > 0 ? 1. : (w < 0 ? -1. : 0.));
        // w = vab . e
        f64 vab_b[3] = {w_b * e[0], w_b * e[1], w_b * e[2]};
        for (int k = 0; k < 3; k++) {
            e_b[k] += w_b * vab[k];
            gA.v[k] += vab_b[k];
            gB.v[k] -= vab_b[k];
        }
        // F = dW/dr (d, h)
        f64 d_b = Fa_b * ka.Wrr + Fb_b * kb.Wrr;
        gA.h += Fa_b * ka.Wrh;
        gB.h += Fb_b * kb.Wrh;
        // EOS
        rho_a_b += Pa_b * dPa_r + csa_b * dca_r;
        gA.u += Pa_b * dPa_u + csa_b * dca_u;
        rho_b_b += Pb_b * dPb_r + csb_b * dcb_r;
        gB.u += Pb_b * dPb_u + csb_b * dcb_u;
        if (p.eos_kind == 2) { // LP07: c_s is the side's own input
            gA.cs += Pa_b * dPa_c + csa_b;
            gB.cs += Pb_b * dPb_c + csb_b;
        }
        // rho = m (hfact / h)^3
        gA.h += rho_a_b * (-3 * rho_a / A.h);
        gB.h += rho_b_b * (-3 * rho_b / B.h);
        // e = r / d, d = |r|
        f64 eb_dot_e = e_b[0] * e[0] + e_b[1] * e[1] + e_b[2] * e[2];
        for (int k = 0; k < 3; k++) gr[k] = (e_b[k] - eb_dot_e * e[k]) * inv + d_b * e[k];
    }

} // namespace shamx
