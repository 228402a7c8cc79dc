// SPDX-License-Identifier: CeCILL-2.1
//
// adjoint.hip -- hand-written adjoint (vector-Jacobian products) of one SPH step: density
// (smoothing length through its density equation), Omega, EOS, force (pressure, constant AV,
// conductivity) and the leapfrog.  No reference counterpart (the reference has no adjoint).
//
// The step map differentiated, for a step of fixed dt:
//   predictor   x' = x + dt v + dt^2/2 a,  v' = v + dt a,  u' = u + dt du     (Solver.cpp:390-524)
//   h'          solves sum_b m W(|x'_a - x'_b|, h_a) = rho_h(h_a)                 (implicit)
//   Omega', P', c_s'  from (x', h', u')                     (ComputeOmega.cpp, ComputeEos.cpp)
//   a', du'     = force(x', v', u', h', Omega')                               (UpdateDerivs.cpp)
//   corrector   v'' = v' + dt/2 (a' - a),  u'' = u' + dt/2 (du' - du)      (integrators.cpp:87)
// with state (x, v, u, a, du).  The smoothing length is differentiated through the implicit
// function theorem: dh_a/dx = -(dF_a/dx) / (dF_a/dh_a), dF_a/dh_a = 3 rho_a Omega_a / h_a.
// All pair loops are in gather form over symmetric neighbour rows (the force criterion
// max(h_a, h_b) and the cache criterion are symmetric): particle a adds the derivative of its
// own pair terms and of the mirrored terms of its neighbours, so there are no atomics.  Ghost
// images carry the seeds of their source particles.
#include "adjoint_pair.hpp"
#include "launch.hpp"

#include <vector>

namespace shamx {

    namespace {
        constexpr int ABLK = 128;

        struct Rows1 {
            const u32 *sel, *outi, *cnt, *off, *ids;
        };

        // out[k][s] = in[k][source of the merged particle at sorted position s]
        struct Gather {
            const f64 *in[4];
            f64 *out[4];
            int n;
        };
        __global__ void k_gather_sorted(u32 nm, u32 N, const u32 *__restrict sort_index,
                                        const u32 *__restrict ghost_src, Gather g) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s >= nm) return;
            u32 b   = sort_index[s];
            u32 src = b < N ? b : ghost_src[b - N];
            for (int k = 0; k < g.n; k++) g.out[k][s] = g.in[k][src];
        }

        // force-record layout (FRec in sph.hip); Omega is read from the merged array
        struct alignas(16) FRecView {
            f64 x, y, z, h, vx, vy, vz, u, P, cs, rho, inv_sub, romega, hR2, alpha, rh, rh4,
                rromega;
        };
        static_assert(sizeof(FRecView) == 144, "force record layout");

        struct ForceGrad {
            f64 *gx, *gy, *gz, *gvx, *gvy, *gvz, *gu, *gh, *gom;
        };

        // a sum in the scalar T: plain for FP64 (the order of the FP64 sums is kept), Kahan
        // compensated for FP32; value() in FP64
        template<class T>
        struct Acc;
        template<>
        struct Acc<f64> {
            f64 s = 0;
            __device__ void add(f64 x) { s += x; }
            __device__ f64 value() const { return s; }
        };
        template<>
        struct Acc<f32> {
            f32 s = 0, c = 0;
            __device__ void add(f32 x) {
                f32 y = x - c, t = s + y;
                c = (t - s) - y;
                s = t;
            }
            __device__ f64 value() const { return (f64) s - (f64) c; }
        };

        // T = f64: every operation FP64.  T = f32 (the FP32 adjoint): r_ab and the neighbour
        // criteria in FP64, the pair terms and their derivatives in FP32, Kahan-compensated
        // sums, results stored in FP64.
        template<class K, class T>
        __global__ void __launch_bounds__(ABLK)
            k_force_vjp(u32 n, Rows1 r, const FRecView *__restrict rec,
                        const f64 *__restrict omega_m, const f64 *__restrict sax,
                        const f64 *__restrict say, const f64 *__restrict saz,
                        const f64 *__restrict sdu, PairParams pp, f64 lp_q, int accum,
                        ForceGrad o) {
            u32 j = blockIdx.x * ABLK + threadIdx.x;
            if (j >= n) return;
            u32 row = r.sel[j], a = r.outi[row];
            const FRecView A = rec[row];
            const PairSideT<T> SA{{T(A.vx), T(A.vy), T(A.vz)}, T(A.u), T(A.h), T(omega_m[a]),
                                  T(A.cs)};
            const T seed_a[3] = {T(sax[row]), T(say[row]), T(saz[row])};
            const T sdu_a     = T(sdu[row]);
            Acc<T> gx[3], gv[3], gu, gh, gom, gcs;
            u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
            for (u32 k = k0; k < k1; k++) {
                u32 b = r.ids[k];
                if (b == row) continue;
                const FRecView B = rec[b];
                const PairSideT<T> SB{{T(B.vx), T(B.vy), T(B.vz)}, T(B.u), T(B.h),
                                      T(omega_m[r.outi[b]]), T(B.cs)};
                T rab[3] = {T(A.x - B.x), T(A.y - B.y), T(A.z - B.z)};
                T acc[3], du, gr[3];
                PairGradT<T> g1, g2;
                // own term: a's outputs seeded by a's seeds
                force_pair<K, T>(pp, rab, SA, SB, acc, du, seed_a, sdu_a, gr, g1, g2);
                for (int c = 0; c < 3; c++) {
                    gx[c].add(gr[c]);
                    gv[c].add(g1.v[c]);
                }
                gu.add(g1.u);
                gh.add(g1.h);
                gom.add(g1.omega);
                gcs.add(g1.cs);
                // mirrored term: b's outputs (pair b, a) seeded by b's seeds; a is side B
                const T seed_b[3] = {T(sax[b]), T(say[b]), T(saz[b])};
                T rba[3]          = {-rab[0], -rab[1], -rab[2]};
                force_pair<K, T>(pp, rba, SB, SA, acc, du, seed_b, T(sdu[b]), gr, g1, g2);
                for (int c = 0; c < 3; c++) {
                    gx[c].add(-gr[c]);
                    gv[c].add(g2.v[c]);
                }
                gu.add(g2.u);
                gh.add(g2.h);
                gom.add(g2.omega);
                gcs.add(g2.cs);
            }
            f64 px = gx[0].value(), py = gx[1].value(), pz = gx[2].value();
            if (pp.eos_kind == 2) {
                // LP07: c_s = c_s0 (|x| / r0)^(-q), so d c_s / dx = -q c_s x / |x|^2
                const f64 R2 = A.x * A.x + A.y * A.y + A.z * A.z;
                const f64 c  = R2 > 0 ? gcs.value() * (-lp_q * A.cs / R2) : 0.;
                px += c * A.x;
                py += c * A.y;
                pz += c * A.z;
            }
            // position, h and Omega gradients add up over the corrector passes of a step (the
            // passes share x, h and Omega); the velocity and u gradients are the pass's own
            if (accum) {
                o.gx[a] += px;
                o.gy[a] += py;
                o.gz[a] += pz;
                o.gh[a] += gh.value();
                o.gom[a] += gom.value();
            } else {
                o.gx[a]  = px;
                o.gy[a]  = py;
                o.gz[a]  = pz;
                o.gh[a]  = gh.value();
                o.gom[a] = gom.value();
            }
            o.gvx[a] = gv[0].value();
            o.gvy[a] = gv[1].value();
            o.gvz[a] = gv[2].value();
            o.gu[a]  = gu.value();
        }

        // Omega_a = 1 + C_a sum_b m dW/dh(d_ab, h_a), C = h / (3 rho_h(h)) = h^4 / (3 m hfact^3),
        // over d_ab <= R h_a: x-gradient (own + mirrored) added to g, h-gradient to gh
        template<class K, class T>
        __global__ void __launch_bounds__(ABLK)
            k_omega_vjp(u32 n, Rows1 r, const double4 *__restrict ps,
                        const f64 *__restrict s_omb, f64 mass, f64 hfact, f64 *__restrict gx,
                        f64 *__restrict gy, f64 *__restrict gz, f64 *__restrict gh) {
            u32 j = blockIdx.x * ABLK + threadIdx.x;
            if (j >= n) return;
            const f64 R = K::Rkern;
            u32 row = r.sel[j], a = r.outi[row];
            double4 A = ps[row];
            f64 h_a   = A.w;
            f64 rho_a = rho_h(mass, h_a, hfact);
            f64 C_a   = h_a / (3 * rho_a);
            f64 ob_a  = s_omb[row];
            const T mT = T(mass), own_c = T(ob_a * C_a * mass);
            Acc<T> S1, S2, g[3];
            u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
            for (u32 k = k0; k < k1; k++) {
                u32 b     = r.ids[k];
                double4 B = ps[b];
                f64 rr[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
                f64 d     = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
                f64 inv   = d > 1e-300 ? 1. / d : 0.;
                if (d <= R * h_a) {
                    KernDeriv<K, T> kd((T) d, (T) h_a);
                    S1.add(mT * kd.Wh);
                    S2.add(mT * kd.Whh);
                    T c = own_c * kd.Wrh * T(inv);
                    for (int t = 0; t < 3; t++) g[t].add(c * T(rr[t]));
                }
                f64 h_b = B.w;
                if (b != row && d <= R * h_b) {
                    KernDeriv<K, T> kd((T) d, (T) h_b);
                    f64 C_b = h_b / (3 * rho_h(mass, h_b, hfact));
                    T c     = T(s_omb[b] * C_b * mass) * kd.Wrh * T(inv);
                    for (int t = 0; t < 3; t++) g[t].add(c * T(rr[t]));
                }
            }
            gx[a] += g[0].value();
            gy[a] += g[1].value();
            gz[a] += g[2].value();
            // dOmega/dh = (4 C / h) S1 + C S2
            gh[a] += ob_a * ((4 * C_a / h_a) * S1.value() + C_a * S2.value());
        }

        // x-gradient of sum_a lambda_a F_a, F_a = sum_b m W(d_ab, h_a) - rho_h(h_a)
        template<class K, class T>
        __global__ void __launch_bounds__(ABLK)
            k_ift_vjp(u32 n, Rows1 r, const double4 *__restrict ps, const f64 *__restrict s_lam,
                      f64 mass, f64 *__restrict gx, f64 *__restrict gy, f64 *__restrict gz) {
            u32 j = blockIdx.x * ABLK + threadIdx.x;
            if (j >= n) return;
            const f64 R = K::Rkern;
            u32 row = r.sel[j], a = r.outi[row];
            double4 A = ps[row];
            f64 lam_a = s_lam[row];
            const T own_c = T(lam_a * mass);
            Acc<T> g[3];
            u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
            for (u32 k = k0; k < k1; k++) {
                u32 b = r.ids[k];
                if (b == row) continue;
                double4 B = ps[b];
                f64 rr[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
                f64 d     = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
                if (!(d > 0)) continue;
                T c = 0;
                if (d <= R * A.w) c += own_c * KernDeriv<K, T>(T(d), T(A.w)).Wr;
                if (d <= R * B.w) c += T(s_lam[b] * mass) * KernDeriv<K, T>(T(d), T(B.w)).Wr;
                c /= T(d);
                for (int t = 0; t < 3; t++) g[t].add(c * T(rr[t]));
            }
            gx[a] += g[0].value();
            gy[a] += g[1].value();
            gz[a] += g[2].value();
        }

        // lambda_a = -hbar_a / (3 rho_a Omega_a / h_a)
        __global__ void k_lambda(u32 n, const f64 *__restrict h, const f64 *__restrict omega,
                                 const f64 *__restrict hbar, f64 mass, f64 hfact,
                                 f64 *__restrict lam) {
            u32 a = blockIdx.x * blockDim.x + threadIdx.x;
            if (a >= n) return;
            f64 rho = rho_h(mass, h[a], hfact);
            lam[a]  = -hbar[a] / (3 * rho * omega[a] / h[a]);
        }

        // hbar_a = rho_bar_a * d rho_h / dh = -3 rho_a / h_a * rho_bar_a
        __global__ void k_rho_to_h(u32 n, const f64 *__restrict h, const f64 *__restrict rbar,
                                   f64 mass, f64 hfact, f64 *__restrict hbar) {
            u32 a = blockIdx.x * blockDim.x + threadIdx.x;
            if (a >= n) return;
            hbar[a] = rbar[a] * (-3 * rho_h(mass, h[a], hfact) / h[a]);
        }

        // corrector adjoint: force seeds and the a/du seeds of the step start
        struct StepSeeds {
            f64 *x[3], *v[3], *u, *a[3], *du;
        };
        __global__ void k_corrector_vjp(u32 n, f64 hdt, StepSeeds s, f64 *__restrict fax,
                                        f64 *__restrict fay, f64 *__restrict faz,
                                        f64 *__restrict fdu) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            f64 *fa[3] = {fax, fay, faz};
            for (int c = 0; c < 3; c++) {
                f64 vb       = s.v[c][i];
                fa[c][i]     = s.a[c][i] + hdt * vb; // seed of a' (= a_{n+1} and in v'')
                s.a[c][i]    = -hdt * vb;            // partial seed of a_n
            }
            f64 ub   = s.u[i];
            fdu[i]   = s.du[i] + hdt * ub;
            s.du[i]  = -hdt * ub;
        }

        // one corrector pass further back: pass i computed a_i = F(v_{i-1}, u_{i-1}) and
        // v_i = v_{i-1} + dt/2 (a_i - a_{i-1}).  With the pass's own v and u gradients g, the
        // seed of v_{i-1} is vbar_i + g, the seed of a_{i-1} (the previous pass's force) is
        // dt/2 g, and the seed of the step start's a is -dt/2 vbar_1 (rewritten at each pass,
        // the last write is pass 2's)
        __global__ void k_pass_vjp(u32 n, f64 hdt, StepSeeds s, ForceGrad g,
                                   f64 *__restrict fax, f64 *__restrict fay,
                                   f64 *__restrict faz, f64 *__restrict fdu) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            f64 *fa[3]       = {fax, fay, faz};
            const f64 gv[3]  = {g.gvx[i], g.gvy[i], g.gvz[i]};
            for (int c = 0; c < 3; c++) {
                f64 vb    = s.v[c][i] + gv[c];
                s.v[c][i] = vb;
                fa[c][i]  = hdt * gv[c];
                s.a[c][i] = -hdt * vb;
            }
            f64 gu  = g.gu[i];
            f64 ub  = s.u[i] + gu;
            s.u[i]  = ub;
            fdu[i]  = hdt * gu;
            s.du[i] = -hdt * ub;
        }

        __global__ void k_accumulate3(u32 n, const f64 *__restrict ax, const f64 *__restrict ay,
                                      const f64 *__restrict az, f64 *__restrict sum) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            sum[i] += ax[i];
            sum[n + i] += ay[i];
            sum[2 * (size_t) n + i] += az[i];
        }

        // add the force / Omega / IFT gradients, then the predictor adjoint
        // Sink-gas gravity (SinkParticlesUpdate.cpp:69-150, ExternalForces.cpp:48-70) reversed
        // for one sink: a_ext,i = -G M r_i / s_i^3 and the sink's a_gas = sum_i G m r_i / s_i^3,
        // r_i = x_i - p, s_i = sqrt(|r_i|^2 + eps^2) (|r_i| without softening), both zero inside
        // the accretion radius.  With J_i = d(r / s^3)/dr = I / s^3 - 3 r r^T / s^5 (symmetric):
        // xbar_i += J_i (G m abar_gas - G M abar_i) and pbar -= the sum over i of the same (per
        // block into part[3 * block + k]).
        struct SinkVjp {
            f64 px, py, pz, M, racc, G, m, soft, bx, by, bz;
        };
        __global__ void __launch_bounds__(256)
            k_sink_gas_vjp(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                           const f64 *__restrict z, SinkVjp k, const f64 *__restrict fax,
                           const f64 *__restrict fay, const f64 *__restrict faz,
                           f64 *__restrict gx, f64 *__restrict gy, f64 *__restrict gz,
                           f64 *__restrict part) {
            __shared__ f64 lds[3][256 / WAVE];
            const u32 i = blockIdx.x * 256 + threadIdx.x;
            f64 c[3]    = {0, 0, 0};
            if (i < n) {
                const f64 r[3] = {x[i] - k.px, y[i] - k.py, z[i] - k.pz};
                const f64 d2   = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (!(sqrt(d2) < k.racc)) {
                    const f64 s2 = k.soft == 0 ? d2 : d2 + k.soft * k.soft;
                    const f64 sv = sqrt(s2), s3 = sv * s2, s5 = s3 * s2;
                    const f64 w[3] = {k.G * k.m * k.bx - k.G * k.M * fax[i],
                                      k.G * k.m * k.by - k.G * k.M * fay[i],
                                      k.G * k.m * k.bz - k.G * k.M * faz[i]};
                    const f64 rw = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
                    for (int t = 0; t < 3; t++) c[t] = w[t] / s3 - 3 * r[t] * rw / s5;
                    gx[i] += c[0];
                    gy[i] += c[1];
                    gz[i] += c[2];
                }
            }
            for (int t = 0; t < 3; t++) {
                f64 v = wave_reduce(c[t], [](f64 p, f64 q) { return p + q; });
                if (__lane_id() == 0) lds[t][threadIdx.x / WAVE] = v;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                f64 v = 0;
                for (int wv = 0; wv < 256 / WAVE; wv++) v += lds[threadIdx.x][wv];
                part[3 * (size_t) blockIdx.x + threadIdx.x] = -v;
            }
        }

        __global__ void k_predictor_vjp(u32 n, f64 dt, StepSeeds s, ForceGrad g) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            f64 gxv[3] = {g.gx[i], g.gy[i], g.gz[i]}, gvv[3] = {g.gvx[i], g.gvy[i], g.gvz[i]};
            for (int c = 0; c < 3; c++) {
                f64 xb = s.x[c][i] + gxv[c]; // seed of x'
                f64 vb = s.v[c][i] + gvv[c]; // seed of v'
                s.x[c][i] = xb;
                s.v[c][i] = vb + dt * xb;
                s.a[c][i] += dt * vb + (dt * (dt * 0.5)) * xb;
            }
            f64 ub  = s.u[i] + g.gu[i];
            s.u[i]  = ub;
            s.du[i] += dt * ub;
        }

        template<class F>
        void dispatch_adjoint_kernel(int id, F &&f) {
            switch (id) {
            case 0: f(M4{}); break;
            case 1: f(M6{}); break;
            case 2: f(M8{}); break;
            case 3: f(C2{}); break;
            default:
                throw std::invalid_argument("adjoint: available for the M4, M6, M8 and C2 kernels");
            }
        }
        // kernel and scalar of the pair terms (FP64, or FP32 when A.fp32)
        template<class F>
        void dispatch_adjoint_scalar(const AdjointArgs &A, F &&f) {
            dispatch_adjoint_kernel(A.kernel, [&](auto K) {
                if (A.fp32) f(K, f32{});
                else f(K, f64{});
            });
        }
    } // namespace

    void adjoint_gather_sorted(hipStream_t s, u32 nm, u32 N, const u32 *sort_index,
                               const u32 *ghost_src, int nf, const f64 *const *in,
                               f64 *const *out) {
        if (nm == 0) return;
        Gather g{};
        g.n = nf;
        for (int k = 0; k < nf; k++) {
            g.in[k]  = in[k];
            g.out[k] = out[k];
        }
        k_gather_sorted<<<div_up(nm, 256), 256, 0, s>>>(nm, N, sort_index, ghost_src, g);
        SHAMX_HIP(hipGetLastError());
    }

    void adjoint_step(hipStream_t s, const AdjointArgs &A, f64 *const *seeds,
                      DevArray<f64> &work, const SinkAdjoint *sk, u32 passes,
                      const std::function<void(u32)> *records_for_pass) {
        const u32 N = A.n, nm = A.nm;
        if (N == 0) return;
        if (passes == 0) throw std::invalid_argument("adjoint: a step has at least one pass");
        const bool sum_ext = sk && sk->count && passes > 1;
        // work: 4 force seeds (local) + 4 sorted force seeds + 9 gradients + 2 sorted (Omega
        // bar, lambda) + lambda (local) [+ 3 sums of the force seeds over the passes]
        size_t need = (size_t) 4 * N + 4 * (size_t) nm + 9 * (size_t) N + 2 * (size_t) nm + N +
                      (sum_ext ? 3 * (size_t) N : 0);
        work.resize(need);
        f64 *w    = work.data();
        f64 *fa[4] = {w, w + N, w + 2 * (size_t) N, w + 3 * (size_t) N};
        w += 4 * (size_t) N;
        f64 *sf[4] = {w, w + nm, w + 2 * (size_t) nm, w + 3 * (size_t) nm};
        w += 4 * (size_t) nm;
        ForceGrad g{w,         w + N,     w + 2 * N, w + 3 * N, w + 4 * N,
                    w + 5 * N, w + 6 * N, w + 7 * N, w + 8 * N};
        w += 9 * (size_t) N;
        f64 *s_omb = w, *s_lam = w + nm;
        w += 2 * (size_t) nm;
        f64 *lam = w;
        w += N;
        f64 *fext = sum_ext ? w : nullptr;

        StepSeeds ss{{seeds[0], seeds[1], seeds[2]},
                     {seeds[3], seeds[4], seeds[5]},
                     seeds[6],
                     {seeds[7], seeds[8], seeds[9]},
                     seeds[10]};
        const f64 dt = A.dt;
        k_corrector_vjp<<<div_up(N, 256), 256, 0, s>>>(N, dt / 2, ss, fa[0], fa[1], fa[2], fa[3]);
        if (sum_ext) fill_f64(s, fext, 0., 3 * (size_t) N);
        PairParams pp{A.mass,     A.hfact,   A.gamma,    A.alpha_u, A.alpha_AV,
                      A.beta_AV,  A.eos_kind, A.cs_iso};
        Rows1 rf{A.lpos, A.sort_index, A.fcnt, A.offset, A.fids};
        Rows1 rc{A.lpos, A.sort_index, A.cnt, A.offset, A.ids};
        dispatch_adjoint_scalar(A, [&](auto K, auto T) {
            using KT = decltype(K);
            using TT = decltype(T);
            // the corrector passes from the last to the first, each on its own records
            for (u32 i = passes; i >= 1; i--) {
                if (records_for_pass) (*records_for_pass)(i);
                if (sum_ext)
                    k_accumulate3<<<div_up(N, 256), 256, 0, s>>>(N, fa[0], fa[1], fa[2], fext);
                const f64 *fin[4] = {fa[0], fa[1], fa[2], fa[3]};
                adjoint_gather_sorted(s, nm, N, A.sort_index, A.ghost_src, 4, fin, sf);
                k_force_vjp<KT, TT><<<div_up(N, ABLK), ABLK, 0, s>>>(
                    N, rf, (const FRecView *) A.frec, A.omega, sf[0], sf[1], sf[2], sf[3], pp,
                    A.lp_q, i < passes ? 1 : 0, g);
                if (i > 1)
                    k_pass_vjp<<<div_up(N, 256), 256, 0, s>>>(N, dt / 2, ss, g, fa[0], fa[1],
                                                                fa[2], fa[3]);
            }
            // Omega bar of the merged particles, then the Omega adjoint
            const f64 *oin[1] = {g.gom};
            f64 *oout[1]      = {s_omb};
            adjoint_gather_sorted(s, nm, N, A.sort_index, A.ghost_src, 1, oin, oout);
            k_omega_vjp<KT, TT><<<div_up(N, ABLK), ABLK, 0, s>>>(
                N, rc, A.ps, s_omb, A.mass, A.hfact, g.gx, g.gy, g.gz, g.gh);
            // smoothing length through its density equation
            k_lambda<<<div_up(N, 256), 256, 0, s>>>(N, A.h, A.omega, g.gh, A.mass, A.hfact,
                                                    lam);
            const f64 *lin[1] = {lam};
            f64 *lout[1]      = {s_lam};
            adjoint_gather_sorted(s, nm, N, A.sort_index, A.ghost_src, 1, lin, lout);
            k_ift_vjp<KT, TT><<<div_up(N, ABLK), ABLK, 0, s>>>(N, rc, A.ps, s_lam, A.mass,
                                                               g.gx, g.gy, g.gz);
        });
        if (sk && sk->count) {
            // the force seeds of the local particles (summed over the passes) are also those of
            // their external accelerations (a = hydro + a_ext, a_ext computed once per step);
            // the gas gradients add to the positions'
            const f64 *ex = sum_ext ? fext : fa[0];
            const f64 *ey = sum_ext ? fext + N : fa[1];
            const f64 *ez = sum_ext ? fext + 2 * (size_t) N : fa[2];
            const u32 nb = div_up(N, 256u);
            DevArray<f64> part(3 * (size_t) nb);
            std::vector<f64> hp(3 * (size_t) nb);
            for (u32 k = 0; k < sk->count; k++) {
                SinkVjp a{sk->pos[3 * k],  sk->pos[3 * k + 1],      sk->pos[3 * k + 2],
                          sk->mass[k],     sk->racc[k],             sk->G,
                          A.mass,          sk->soft,                sk->abar_gas[3 * k],
                          sk->abar_gas[3 * k + 1], sk->abar_gas[3 * k + 2]};
                k_sink_gas_vjp<<<nb, 256, 0, s>>>(N, sk->x, sk->y, sk->z, a, ex, ey, ez, g.gx,
                                                  g.gy, g.gz, part.data());
                SHAMX_HIP(hipMemcpyAsync(hp.data(), part.data(), hp.size() * sizeof(f64),
                                         hipMemcpyDeviceToHost, s));
                SHAMX_HIP(hipStreamSynchronize(s));
                for (int t = 0; t < 3; t++) {
                    f64 acc = 0;
                    for (u32 b = 0; b < nb; b++) acc += hp[3 * (size_t) b + t];
                    sk->pbar[3 * k + t] = acc;
                }
            }
        }
        k_predictor_vjp<<<div_up(N, 256), 256, 0, s>>>(N, dt, ss, g);
        SHAMX_HIP(hipGetLastError());
    }

    void adjoint_density(hipStream_t s, const AdjointArgs &A, const f64 *rho_bar, f64 *gx,
                         f64 *gy, f64 *gz, DevArray<f64> &work) {
        const u32 N = A.n, nm = A.nm;
        if (N == 0) return;
        work.resize((size_t) 2 * N + nm);
        f64 *hbar = work.data(), *lam = hbar + N, *s_lam = lam + N;
        k_rho_to_h<<<div_up(N, 256), 256, 0, s>>>(N, A.h, rho_bar, A.mass, A.hfact, hbar);
        k_lambda<<<div_up(N, 256), 256, 0, s>>>(N, A.h, A.omega, hbar, A.mass, A.hfact, lam);
        const f64 *lin[1] = {lam};
        f64 *lout[1]      = {s_lam};
        adjoint_gather_sorted(s, nm, N, A.sort_index, A.ghost_src, 1, lin, lout);
        Rows1 rc{A.lpos, A.sort_index, A.cnt, A.offset, A.ids};
        dispatch_adjoint_scalar(A, [&](auto K, auto T) {
            k_ift_vjp<decltype(K), decltype(T)><<<div_up(N, ABLK), ABLK, 0, s>>>(
                N, rc, A.ps, s_lam, A.mass, gx, gy, gz);
        });
        SHAMX_HIP(hipGetLastError());
    }

} // namespace shamx
