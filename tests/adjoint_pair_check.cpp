// Host build of the adjoint pair header (test harness): the force pair's forward value and
// its hand-written reverse sweep, so tests compare the gradient with finite differences.
#include "../shamrock_amd/csrc/adjoint_pair.hpp"

using namespace shamx;

namespace {
    template<class K>
    void run(const PairParams &p, const double *in, const double *seeds, double *loss,
             double *grad) {
        const double r[3] = {in[0], in[1], in[2]};
        PairSide A{{in[3], in[4], in[5]}, in[9], in[11], in[13], in[15]};
        PairSide B{{in[6], in[7], in[8]}, in[10], in[12], in[14], in[16]};
        double acc[3], du, gr[3];
        PairGrad gA{}, gB{};
        force_pair<K>(p, r, A, B, acc, du, grad ? seeds : nullptr, seeds[3], gr, gA, gB);
        *loss = seeds[0] * acc[0] + seeds[1] * acc[1] + seeds[2] * acc[2] + seeds[3] * du;
        if (!grad) return;
        double g[17] = {gr[0], gr[1], gr[2], gA.v[0], gA.v[1], gA.v[2], gB.v[0], gB.v[1],
                        gB.v[2], gA.u, gB.u, gA.h, gB.h, gA.omega, gB.omega, gA.cs, gB.cs};
        for (int k = 0; k < 17; k++) grad[k] = g[k];
    }
} // namespace

// in[17] = r(3), va(3), vb(3), ua, ub, ha, hb, omega_a, omega_b, cs_a, cs_b (LP07 sound
// speeds); seeds[4] = (a_x, a_y, a_z, du)
// par[8] = mass, hfact, gamma, alpha_u, alpha_AV, beta_AV, eos_kind, cs_iso
extern "C" int pair_eval(int kernel, const double *par, const double *in, const double *seeds,
                         double *loss, double *grad) {
    PairParams p{par[0], par[1], par[2], par[3], par[4], par[5], (int) par[6], par[7]};
    switch (kernel) {
    case 0: run<M4>(p, in, seeds, loss, grad); return 0;
    case 1: run<M6>(p, in, seeds, loss, grad); return 0;
    case 2: run<M8>(p, in, seeds, loss, grad); return 0;
    case 3: run<C2>(p, in, seeds, loss, grad); return 0;
    default: return 1;
    }
}

// W and its partial derivatives (r, h): out[6] = W, Wr, Wh, Wrr, Wrh, Whh
extern "C" int kern_deriv(int kernel, double r, double h, double *out) {
    auto put = [&](auto k) {
        out[0] = k.W, out[1] = k.Wr, out[2] = k.Wh, out[3] = k.Wrr, out[4] = k.Wrh,
        out[5] = k.Whh;
    };
    switch (kernel) {
    case 0: put(KernDeriv<M4>(r, h)); return 0;
    case 1: put(KernDeriv<M6>(r, h)); return 0;
    case 2: put(KernDeriv<M8>(r, h)); return 0;
    case 3: put(KernDeriv<C2>(r, h)); return 0;
    default: return 1;
    }
}
