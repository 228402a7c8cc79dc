// SPDX-License-Identifier: CeCILL-2.1
//
// common.hpp -- shared device/host definitions of the MI355X SPH step library.
//
// Floating point convention: the whole library is compiled with -ffp-contract=off, so every
// + - * / in FP64 is a single IEEE operation in the order written.  The arithmetic below keeps
// the operation order of the reference expressions it cites, which makes the FP64 path
// reproduce the reference step bit for bit where the reference itself is deterministic
// (neighbour criteria, Morton codes, tree boxes) and to the last few ulp elsewhere.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace shamx {

    using u8  = std::uint8_t;
    using u16 = std::uint16_t;
    using u32 = std::uint32_t;
    using u64 = std::uint64_t;
    using i32 = std::int32_t;
    using i64 = std::int64_t;
    using f32 = float;
    using f64 = double;

    struct HipError : std::runtime_error {
        using std::runtime_error::runtime_error;
    };

#define SHAMX_HIP(call)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::shamx::HipError(std::string(#call " failed: ") + hipGetErrorString(e_) +      \
                                    " at " __FILE__ ":" + std::to_string(__LINE__));              \
    } while (0)

    inline u32 div_up(u64 a, u64 b) { return (u32) ((a + b - 1) / b); }

    constexpr int WAVE = 64; // CDNA wavefront width

    // ---------------------------------------------------------------------------------------
    // SPH kernels (reference: shammath/include/shammath/sphkernels.hpp; M4 :29-117,
    // M6 :265-405, M8 :584-680, C2 :1181-1262, C4 :1263-1359, C6 :1360-1460, generic
    // W_3d / dW_3d / dhW_3d :2329-2343).  pi = 3.141592653589793116 (constants.hpp:53).
    // Templated on the scalar so the mixed-precision force loop reuses the same polynomials.
    // ---------------------------------------------------------------------------------------
    constexpr f64 PI = 3.141592653589793116;

    template<class T>
    __host__ __device__ inline T p2(T a) { return a * a; }

    struct M4 {
        static constexpr f64 Rkern = 2, hfactd = 1.2, norm_3d = 1 / PI;
        static constexpr int id = 0;
        template<class T>
        __host__ __device__ static T f(T q) {
            T t1 = T(2) - q, t2 = T(1) - q;
            t1 = t1 * t1 * t1;
            t2 = t2 * t2 * t2;
            t1 *= T(1. / 4.);
            t2 *= T(-1);
            if (q < T(1)) return t1 + t2;
            if (q < T(2)) return t1;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            if (q < T(1)) return T(-3) * q + T(9. / 4.) * q * q;
            if (q < T(2)) return T(-3) + T(3) * q - T(3. / 4.) * q * q;
            return T(0);
        }
    };
    struct M6 {
        static constexpr f64 Rkern = 3, hfactd = 1.0, norm_3d = 1 / (120 * PI);
        static constexpr int id = 1;
        template<class T>
        __host__ __device__ static T f(T q) {
            T t1 = T(3) - q, t2 = T(2) - q, t3 = T(1) - q;
            T t1_2 = t1 * t1, t2_2 = t2 * t2, t3_2 = t3 * t3;
            t1 = t1 * t1_2 * t1_2;
            t2 = t2 * t2_2 * t2_2;
            t3 = t3 * t3_2 * t3_2;
            t2 *= T(-6);
            t3 *= T(15);
            if (q < T(1)) return t1 + t2 + t3;
            if (q < T(2)) return t1 + t2;
            if (q < T(3)) return t1;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            T t1 = T(3) - q, t2 = T(2) - q, t3 = T(1) - q;
            T t1_2 = t1 * t1, t2_2 = t2 * t2, t3_2 = t3 * t3;
            t1 = t1_2 * t1_2;
            t2 = t2_2 * t2_2;
            t3 = t3_2 * t3_2;
            t1 *= T(-5);
            t2 *= T(30);
            t3 *= T(-75);
            if (q < T(1)) return t1 + t2 + t3;
            if (q < T(2)) return t1 + t2;
            if (q < T(3)) return t1;
            return T(0);
        }
    };
    template<class T>
    __host__ __device__ inline T pow6(T a) {
        T t = a * a * a; // pow_constexpr<3> = (a*a)*a
        return t * t;
    }
    template<class T>
    __host__ __device__ inline T pow7(T a) {
        T t = a * a * a; // pow_constexpr<7> = (p3*p3)*a
        return t * t * a;
    }
    struct M8 {
        static constexpr f64 Rkern = 4.0, hfactd = 1.0, norm_3d = (1.0 / 6720.0) / PI;
        static constexpr int id = 2;
        template<class T>
        __host__ __device__ static T f(T q) {
            T t1 = pow7(T(4) - q), t2 = pow7(T(3) - q), t3 = pow7(T(2) - q), t4 = pow7(T(1) - q);
            t2 *= T(-8.);
            t3 *= T(28.);
            t4 *= T(-56.);
            if (q < T(1)) return t1 + t2 + t3 + t4;
            if (q < T(2)) return t1 + t2 + t3;
            if (q < T(3)) return t1 + t2;
            if (q < T(4)) return t1;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            T t1 = pow6(T(4) - q), t2 = pow6(T(3) - q), t3 = pow6(T(2) - q), t4 = pow6(T(1) - q);
            t1 *= T(-7.);
            t2 *= T(56.);
            t3 *= T(-196.);
            t4 *= T(392.);
            if (q < T(1)) return t1 + t2 + t3 + t4;
            if (q < T(2)) return t1 + t2 + t3;
            if (q < T(3)) return t1 + t2;
            if (q < T(4)) return t1;
            return T(0);
        }
    };
    struct C2 {
        static constexpr f64 Rkern = 2, hfactd = 1.4, norm_3d = 21 / (16 * PI);
        static constexpr int id = 3;
        template<class T>
        __host__ __device__ static T f(T q) {
            T p1 = T(1) - q * T(0.5), pq = T(1) + q * T(2);
            p1 *= p1;
            p1 *= p1;
            if (q < T(2)) return p1 * pq;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            T p1 = T(1) - q * T(0.5), pq = T(1) + q * T(2);
            T p3 = p1 * p1 * p1;
            if (q < T(2)) return T(2) * (p3 * p1 - p3 * pq);
            return T(0);
        }
    };
    struct C4 {
        static constexpr f64 Rkern = 2, hfactd = 1.6, norm_3d = 495. / (256. * PI);
        static constexpr int id = 4;
        template<class T>
        __host__ __device__ static T f(T q) {
            T p1 = T(1) - q * T(0.5), pq = T(1) + q * T(3) + T(35. / 12.) * q * q;
            p1 *= p1;
            p1 = p1 * p1 * p1;
            if (q < T(2)) return p1 * pq;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            T p1 = T(-2) + q, pq = T(2) + T(5) * q;
            T p14 = p1 * p1;
            p14 *= p14;
            if (q < T(2)) return T(7. / 96.) * p14 * p1 * q * pq;
            return T(0);
        }
    };
    struct C6 {
        static constexpr f64 Rkern = 2, hfactd = 2.2, norm_3d = 1365. / (512. * PI);
        static constexpr int id = 5;
        template<class T>
        __host__ __device__ static T f(T q) {
            T p1 = T(1) - q * T(0.5);
            T pq = T(1) + q * T(4) + T(25. / 4.) * q * q + T(4) * q * q * q;
            p1 *= p1;
            p1 *= p1;
            p1 *= p1;
            if (q < T(2)) return p1 * pq;
            return T(0);
        }
        template<class T>
        __host__ __device__ static T df(T q) {
            T p1 = T(-2) + q, pq = T(2) + T(7) * q + T(8) * q * q;
            T p12 = p1 * p1, p14 = p12 * p12, p17 = p12 * p14 * p1;
            if (q < T(2)) return T(11. / 512.) * p17 * q * pq;
            return T(0);
        }
    };

    template<class K, class T>
    struct Kern {
        static constexpr T Rkern = T(K::Rkern), hfactd = T(K::hfactd), norm = T(K::norm_3d);
        __host__ __device__ static T W(T r, T h) { return norm * K::f(r / h) / (h * h * h); }
        __host__ __device__ static T dW(T r, T h) { return norm * K::df(r / h) / (h * h * h * h); }
        __host__ __device__ static T dhW(T r, T h) {
            return -(norm) * (T(3) * K::f(r / h) + (r / h) * K::df(r / h)) / (h * h * h * h);
        }
    };

    // density.hpp:22-41
    template<class T>
    __host__ __device__ inline T rho_h(T m, T h, T hfact) {
        return m * (hfact / h) * (hfact / h) * (hfact / h);
    }

    // kernel id dispatch (the reference instantiates the solver per kernel: Solver.cpp:3276-3282)
    template<class F>
    inline void dispatch_kernel(int id, F &&f) {
        switch (id) {
        case 0: f(M4{}); break;
        case 1: f(M6{}); break;
        case 2: f(M8{}); break;
        case 3: f(C2{}); break;
        case 4: f(C4{}); break;
        case 5: f(C6{}); break;
        default: throw std::invalid_argument("unknown SPH kernel id " + std::to_string(id));
        }
    }

    // ---------------------------------------------------------------------------------------
    // order-preserving maps of doubles onto u64 for atomicMax / atomicMin reductions
    // ---------------------------------------------------------------------------------------
    __device__ inline u64 f64_to_ordered(f64 v) {
        u64 b = __double_as_longlong(v);
        return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
    }
    __host__ __device__ inline f64 ordered_to_f64(u64 k) {
        u64 b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        f64 v;
        __builtin_memcpy(&v, &b, 8);
        return v;
    }
    constexpr u64 ORDERED_NEG_INF = 0x000fffffffffffffull; // below every ordered finite key
    constexpr u64 ORDERED_POS_INF = 0xfff0000000000000ull; // above every ordered finite key

    // wave-level reductions (64 lanes)
    template<class T, class Op>
    __device__ inline T wave_reduce(T v, Op op) {
        for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, WAVE));
        return v;
    }

} // namespace shamx
