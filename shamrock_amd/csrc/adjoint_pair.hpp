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
        template<class T>
        __host__ __device__ static T ddf(T q) {
            if (q < T(1)) return T(-3) + T(4.5) * q;
            if (q < T(2)) return T(3) - T(1.5) * q;
            return T(0);
        }
    };
    template<>
    struct KernD2<M6> {
        template<class T>
        __host__ __device__ static T ddf(T q) {
            T t1 = T(3) - q, t2 = T(2) - q, t3 = T(1) - q;
            T a = T(20) * t1 * t1 * t1, b = T(-120) * t2 * t2 * t2, c = T(300) * t3 * t3 * t3;
            if (q < T(1)) return a + b + c;
            if (q < T(2)) return a + b;
            if (q < T(3)) return a;
            return T(0);
        }
    };
    template<>
    struct KernD2<M8> {
        template<class T>
        __host__ __device__ static T ddf(T q) {
            auto p5 = [](T v) { return v * v * v * v * v; };
            T a = T(42) * p5(T(4) - q), b = T(-336) * p5(T(3) - q), c = T(1176) * p5(T(2) - q),
              d = T(-2352) * p5(T(1) - q);
            if (q < T(1)) return a + b + c + d;
            if (q < T(2)) return a + b + c;
            if (q < T(3)) return a + b;
            if (q < T(4)) return a;
            return T(0);
        }
    };
    template<>
    struct KernD2<C2> {
        template<class T>
        __host__ __device__ static T ddf(T q) {
            T p1 = T(1) - T(0.5) * q;
            if (q < T(2)) return p1 * p1 * (T(10) * q - T(5));
            return T(0);
        }
    };
    template<class K>
    constexpr bool has_adjoint() {
        return K::id == M4::id || K::id == M6::id || K::id == M8::id || K::id == C2::id;
    }

    // W(r, h) and its first and second partial derivatives in r and h, in the scalar T
    template<class K, class T = f64>
    struct KernDeriv {
        T W, Wr, Wh, Wrr, Wrh, Whh;
        __host__ __device__ KernDeriv(T r, T h) {
            const T norm = T(K::norm_3d);
            T q = r / h, f = K::f(q), df = K::df(q), ddf = KernD2<K>::ddf(q);
            T h2 = h * h, h3 = h2 * h, h4 = h3 * h, h5 = h4 * h;
            W   = norm * f / h3;
            Wr  = norm * df / h4;
            Wh  = -norm * (T(3) * f + q * df) / h4;
            Wrr = norm * ddf / h5;
            Wrh = -norm * (T(4) * df + q * ddf) / h5;
            Whh = norm * (T(12) * f + T(8) * q * df + q * q * ddf) / h5;
        }
    };

    struct PairParams {
        f64 mass, hfact, gamma, alpha_u, alpha_AV, beta_AV;
        int eos_kind; // 0 adiabatic, 1 isothermal (sound speed cs_iso), 2 LP07 (side's cs)
        f64 cs_iso;
    };

    // per-particle state entering a force pair (cs: the particle's LP07 sound speed, which
    // depends on its position; unused by the other equations of state)
    template<class T>
    struct PairSideT {
        T v[3], u, h, omega, cs;
    };
    // gradient of the seeded pair outputs w.r.t. one side's inputs
    template<class T>
    struct PairGradT {
        T v[3], u, h, omega, cs;
    };
    using PairSide = PairSideT<f64>;
    using PairGrad = PairGradT<f64>;

    // P, c_s and their partial derivatives; for LP07 (ComputeEos.cpp:726-798) P = c_s^2 rho
    // with c_s the side's own (dP_dcs = 2 c_s rho, c_s passes through)
    template<class T>
    __host__ __device__ inline void eos_of(const PairParams &p, T rho, T u, T cs_loc, T &P,
                                           T &cs, T &dP_drho, T &dP_du, T &dcs_drho, T &dcs_du,
                                           T &dP_dcs) {
        dP_dcs = 0;
        const T gamma = T(p.gamma), cs_iso = T(p.cs_iso);
        if (p.eos_kind == 2) {
            P        = cs_loc * cs_loc * rho;
            cs       = cs_loc;
            dP_drho  = cs_loc * cs_loc;
            dP_du    = 0;
            dcs_drho = 0;
            dcs_du   = 0;
            dP_dcs   = T(2) * cs_loc * rho;
        } else if (p.eos_kind == 0) {
            P  = (gamma - T(1)) * rho * u;
            cs = std::sqrt(gamma * P / rho);
            dP_drho = (gamma - T(1)) * u;
            dP_du   = (gamma - T(1)) * rho;
            // c_s = sqrt(gamma (gamma - 1) u): independent of rho
            dcs_drho = 0;
            dcs_du   = cs > T(0) ? gamma * (gamma - T(1)) / (T(2) * cs) : T(0);
        } else {
            P        = cs_iso * cs_iso * rho;
            cs       = cs_iso;
            dP_drho  = cs_iso * cs_iso;
            dP_du    = 0;
            dcs_drho = 0;
            dcs_du   = 0;
        }
    }

    // Contribution of neighbour b to (a_a, du_a) (forward), or, with seeds (ab[3], db) for
    // (a_a, du_a), the reverse sweep: gradients w.r.t. r = x_a - x_b (gr), side a and side b.
    // T is the scalar of every operation (f64, or f32 for the FP32 adjoint).
    template<class K, class T = f64>
    __host__ __device__ inline void force_pair(const PairParams &p, const T r[3],
                                               const PairSideT<T> &A, const PairSideT<T> &B,
                                               T acc[3], T &du, const T *ab, T db, T gr[3],
                                               PairGradT<T> &gA, PairGradT<T> &gB) {
        const T m = T(p.mass), hfact = T(p.hfact), alpha_AV = T(p.alpha_AV),
                beta_AV = T(p.beta_AV), alpha_u = T(p.alpha_u);
        // ---- forward
        T d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        T d  = std::sqrt(d2);
        T inv = (d >= T(1e-9)) ? T(1) / d : T(0);
        T e[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
        T hfa = hfact / A.h, hfb = hfact / B.h;
        T rho_a = m * hfa * hfa * hfa, rho_b = m * hfb * hfb * hfb;
        T Pa, csa, dPa_r, dPa_u, dca_r, dca_u, dPa_c, Pb, csb, dPb_r, dPb_u, dcb_r, dcb_u, dPb_c;
        eos_of<T>(p, rho_a, A.u, A.cs, Pa, csa, dPa_r, dPa_u, dca_r, dca_u, dPa_c);
        eos_of<T>(p, rho_b, B.u, B.cs, Pb, csb, dPb_r, dPb_u, dcb_r, dcb_u, dPb_c);
        KernDeriv<K, T> ka(d, A.h), kb(d, B.h);
        T Fa = ka.Wr, Fb = kb.Wr;
        T vab[3] = {A.v[0] - B.v[0], A.v[1] - B.v[1], A.v[2] - B.v[2]};
        T w  = vab[0] * e[0] + vab[1] * e[1] + vab[2] * e[2];
        T aw = std::fabs(w);
        T vsa = alpha_AV * csa + beta_AV * aw;
        T vsb = alpha_AV * csb + beta_AV * aw;
        T ravg = (rho_a + rho_b) * T(0.5);
        T dP   = Pa - Pb;
        T vsu  = std::sqrt(std::fabs(dP) / ravg);
        T qa_in = T(-0.5) * rho_a * vsa * w, qb_in = T(-0.5) * rho_b * vsb * w;
        T qa = qa_in > T(0) ? qa_in : T(0), qb = qb_in > T(0) ? qb_in : T(0);
        T APa = Pa + qa, APb = Pb + qb;
        T isa = T(1) / (rho_a * rho_a * A.omega), isb = T(1) / (rho_b * rho_b * B.omega);
        T Sa = APa * isa, Sb = APb * isb;
        T G  = Sa * Fa + Sb * Fb;
        T Ya = Fa / (rho_a * A.omega), Yb = Fb / (rho_b * B.omega);
        T Y  = Ya + Yb;
        T cu = m * alpha_u * T(0.5);
        T dU = A.u - B.u;
        for (int k = 0; k < 3; k++) acc[k] = -m * G * e[k];
        du = m * Sa * Fa * w + cu * vsu * dU * Y;
        if (!ab) return;
        // ---- reverse
        gA = PairGradT<T>{{0, 0, 0}, 0, 0, 0, 0};
        gB = PairGradT<T>{{0, 0, 0}, 0, 0, 0, 0};
        gr[0] = gr[1] = gr[2] = 0;
        if (d < T(1e-9)) return;
        T ae    = ab[0] * e[0] + ab[1] * e[1] + ab[2] * e[2];
        T e_b[3] = {-m * G * ab[0], -m * G * ab[1], -m * G * ab[2]};
        T G_b   = -m * ae;
        // du = m Sa Fa w + cu vsu dU Y
        T Sa_b = db * m * Fa * w, Fa_b = db * m * Sa * w, w_b = db * m * Sa * Fa;
        T vsu_b = db * cu * dU * Y;
        T dU_b  = db * cu * vsu * Y;
        T Y_b   = db * cu * vsu * dU;
        gA.u += dU_b;
        gB.u -= dU_b;
        T rho_a_b = 0, rho_b_b = 0, Fb_b = 0, Sb_b = 0;
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
        T APa_b = Sa_b * isa, APb_b = Sb_b * isb;
        rho_a_b += T(-2) * Sa_b * Sa / rho_a;
        gA.omega += -Sa_b * Sa / A.omega;
        rho_b_b += T(-2) * Sb_b * Sb / rho_b;
        gB.omega += -Sb_b * Sb / B.omega;
        // AP = P + q
        T Pa_b = APa_b, Pb_b = APb_b, qa_b = APa_b, qb_b = APb_b;
        T vsa_b = 0, vsb_b = 0;
        if (qa_in > T(0)) {
            rho_a_b += qa_b * (T(-0.5) * vsa * w);
            vsa_b += qa_b * (T(-0.5) * rho_a * w);
            w_b += qa_b * (T(-0.5) * rho_a * vsa);
        }
        if (qb_in > T(0)) {
            rho_b_b += qb_b * (T(-0.5) * vsb * w);
            vsb_b += qb_b * (T(-0.5) * rho_b * w);
            w_b += qb_b * (T(-0.5) * rho_b * vsb);
        }
        // vsu = sqrt(|dP| / ravg)
        if (vsu > T(0)) {
            T z_b = vsu_b / (T(2) * vsu);
            T s   = dP > T(0) ? T(1) : T(-1);
            T dP_b = z_b * s / ravg;
            Pa_b += dP_b;
            Pb_b -= dP_b;
            T ravg_b = -z_b * std::fabs(dP) / (ravg * ravg);
            rho_a_b += T(0.5) * ravg_b;
            rho_b_b += T(0.5) * ravg_b;
        }
        // vs = alpha cs + beta |w|
        T csa_b = alpha_AV * vsa_b, csb_b = alpha_AV * vsb_b;
        T aw_b  = beta_AV * (vsa_b + vsb_b);
        w_b += aw_b * (w > T(0) ? T(1) : (w < T(0) ? T(-1) : T(0)));
        // w = vab . e
        T vab_b[3] = {w_b * e[0], w_b * e[1], w_b * e[2]};
        for (int k = 0; k < 3; k++) {
            e_b[k] += w_b * vab[k];
            gA.v[k] += vab_b[k];
            gB.v[k] -= vab_b[k];
        }
        // F = dW/dr (d, h)
        T d_b = Fa_b * ka.Wrr + Fb_b * kb.Wrr;
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
        gA.h += rho_a_b * (T(-3) * rho_a / A.h);
        gB.h += rho_b_b * (T(-3) * rho_b / B.h);
        // e = r / d, d = |r|
        T eb_dot_e = e_b[0] * e[0] + e_b[1] * e[1] + e_b[2] * e[2];
        for (int k = 0; k < 3; k++) gr[k] = (e_b[k] - eb_dot_e * e[k]) * inv + d_b * e[k];
    }

} // namespace shamx
