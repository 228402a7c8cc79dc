// SPDX-License-Identifier: CeCILL-2.1
//
// sph_oracle.hpp -- CPU restatement of the reference SPH timestep, used ONLY as a test oracle
// and as the reported CPU baseline.  It is test infrastructure: product code (the HIP library
// under shamrock_amd/) never includes, links or calls anything in this directory.
//
// Every function cites the reference file:line it restates (paths relative to the reference
// source tree `src/`).  Floating-point expressions keep the reference's operation order; the
// library is compiled with -ffp-contract=off -fno-fast-math so each + - * / sqrt is one IEEE
// double operation.  Where the reference's SYCL compiler may contract a*b+c into an FMA (which
// is implementation defined), this restatement does not contract; the HIP product is compiled
// with the same convention, so both sides round identically.
//
// Pinning (see tests/test_oracle_*.py):
//   morton_codes, bitonic_sort_by_key, reduction, karras, tree aabb / field, clbvh traversal:
//     golden vectors transcribed from the reference unit tests (tests/golden/*.json).
//   iterate_h: golden sequences of IterateSmoothingLengthDensityTests.cpp:250-366.
//   SPH kernels: normalisation integrals, f(R)=0, finite-difference df, closed forms.
//   neighbour cache: brute-force O(N^2) symmetric criterion on small inputs.
//   eos / force / corrector / cfl: closed forms, momentum conservation, zero force on a
//     uniform periodic lattice, two-particle hand computation.
//   ghosts (periodic images): brute-force image enumeration.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

    using u8  = std::uint8_t;
    using u32 = std::uint32_t;
    using u64 = std::uint64_t;
    using i32 = std::int32_t;
    using f64 = double;

    struct vec3 {
        f64 x, y, z;
    };
    inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
    inline vec3 operator*(f64 s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
    inline vec3 operator*(vec3 a, f64 s) { return {a.x * s, a.y * s, a.z * s}; }
    inline vec3 operator/(vec3 a, f64 s) { return {a.x / s, a.y / s, a.z / s}; }
    // sycl::dot on a 3-vector: (x*x + y*y) + z*z, left to right
    inline f64 dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    inline vec3 vmin(vec3 a, vec3 b) {
        return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
    }
    inline vec3 vmax(vec3 a, vec3 b) {
        return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
    }

    // ------------------------------------------------------------------------------------
    // SPH kernels: shammath/include/shammath/sphkernels.hpp
    //   M4 :29-117, M6 :265-405, M8 :584-680, C2 :1181-1262, C4 :1263-1359, C6 :1360-1460,
    //   SPHKernelGen::{W_3d,dW_3d,dhW_3d} :2329-2343.  pi = 3.141592653589793116
    //   (shambase/include/shambase/constants.hpp:53)
    // ------------------------------------------------------------------------------------
    constexpr f64 pi = 3.141592653589793116;

    template<int P>
    inline f64 ipow(f64 a) { // sham::pow_constexpr, shambackends/math.hpp:724-739
        if constexpr (P == 0) {
            return 1.;
        } else if constexpr (P == 1) {
            return a;
        } else if constexpr (P % 2 == 0) {
            f64 t = ipow<P / 2>(a);
            return t * t;
        } else {
            f64 t = ipow<(P - 1) / 2>(a);
            return t * t * a;
        }
    }

    struct KM4 {
        static constexpr f64 Rkern = 2, hfactd = 1.2, norm_3d = 1 / pi;
        static f64 f(f64 q) {
            f64 t1 = 2 - q, t2 = 1 - q;
            t1 = t1 * t1 * t1;
            t2 = t2 * t2 * t2;
            constexpr f64 div1_4 = (1. / 4.);
            t1 *= div1_4;
            t2 *= -1;
            if (q < 1) return t1 + t2;
            else if (q < 2) return t1;
            return 0;
        }
        static f64 df(f64 q) {
            constexpr f64 div9_4 = (9. / 4.), div3_4 = (3. / 4.);
            if (q < 1) return -3 * q + div9_4 * q * q;
            else if (q < 2) return -3 + 3 * q - div3_4 * q * q;
            return 0;
        }
    };
    struct KM6 {
        static constexpr f64 Rkern = 3, hfactd = 1.0, norm_3d = 1 / (120 * pi);
        static f64 f(f64 q) {
            f64 t1 = 3 - q, t2 = 2 - q, t3 = 1 - q;
            f64 t1_2 = t1 * t1, t2_2 = t2 * t2, t3_2 = t3 * t3;
            t1 = t1 * t1_2 * t1_2;
            t2 = t2 * t2_2 * t2_2;
            t3 = t3 * t3_2 * t3_2;
            t1 *= 1;
            t2 *= -6;
            t3 *= 15;
            if (q < 1.) return t1 + t2 + t3;
            else if (q < 2.) return t1 + t2;
            else if (q < 3.) return t1;
            return 0;
        }
        static f64 df(f64 q) {
            f64 t1 = 3 - q, t2 = 2 - q, t3 = 1 - q;
            f64 t1_2 = t1 * t1, t2_2 = t2 * t2, t3_2 = t3 * t3;
            t1 = t1_2 * t1_2;
            t2 = t2_2 * t2_2;
            t3 = t3_2 * t3_2;
            t1 *= (1) * (-5);
            t2 *= (-6) * (-5);
            t3 *= (15) * (-5);
            if (q < 1.) return t1 + t2 + t3;
            else if (q < 2.) return t1 + t2;
            else if (q < 3.) return t1;
            return 0;
        }
    };
    struct KM8 {
        static constexpr f64 Rkern = 4.0, hfactd = 1.0, norm_3d = (1.0 / 6720.0) / pi;
        static f64 f(f64 q) {
            f64 t1 = 4 - q, t2 = 3 - q, t3 = 2 - q, t4 = 1 - q;
            t1 = ipow<7>(t1);
            t2 = ipow<7>(t2);
            t3 = ipow<7>(t3);
            t4 = ipow<7>(t4);
            t1 *= 1.;
            t2 *= -8.;
            t3 *= 28.;
            t4 *= -56.;
            if (q < 1) return t1 + t2 + t3 + t4;
            else if (q < 2) return t1 + t2 + t3;
            else if (q < 3) return t1 + t2;
            else if (q < 4) return t1;
            return 0;
        }
        static f64 df(f64 q) {
            f64 t1 = 4 - q, t2 = 3 - q, t3 = 2 - q, t4 = 1 - q;
            t1 = ipow<6>(t1);
            t2 = ipow<6>(t2);
            t3 = ipow<6>(t3);
            t4 = ipow<6>(t4);
            t1 *= (-7.) * (1.);
            t2 *= (-7.) * (-8.);
            t3 *= (-7.) * (28.);
            t4 *= (-7.) * (-56.);
            if (q < 1) return t1 + t2 + t3 + t4;
            else if (q < 2) return t1 + t2 + t3;
            else if (q < 3) return t1 + t2;
            else if (q < 4) return t1;
            return 0;
        }
    };
    struct KC2 {
        static constexpr f64 Rkern = 2, hfactd = 1.4, norm_3d = 21 / (16 * pi);
        static f64 f(f64 q) {
            constexpr f64 div1_2 = (1. / 2.);
            f64 p1 = (1 - q * div1_2), p2 = (1 + q * 2);
            p1 *= p1;
            p1 *= p1;
            if (q < 2.) return p1 * p2;
            return 0;
        }
        static f64 df(f64 q) {
            constexpr f64 div1_2 = (1. / 2.);
            f64 p1 = (1 - q * div1_2), p2 = (1 + q * 2);
            f64 p3 = p1 * p1 * p1;
            if (q < 2.) return 2 * (p3 * p1 - p3 * p2);
            return 0;
        }
    };
    struct KC4 {
        static constexpr f64 Rkern = 2, hfactd = 1.6, norm_3d = 495. / (256. * pi);
        static f64 f(f64 q) {
            constexpr f64 div1_2 = (1. / 2.), div35_12 = (35. / 12.);
            f64 p1 = (1 - q * div1_2), p2 = (1 + q * 3 + div35_12 * q * q);
            p1 *= p1;
            p1 = p1 * p1 * p1;
            if (q < 2.) return p1 * p2;
            return 0;
        }
        static f64 df(f64 q) {
            constexpr f64 div7_96 = (7. / 96.);
            f64 p1 = (-2 + q), p2 = (2 + 5 * q);
            f64 p14 = p1 * p1;
            p14 *= p14;
            if (q < 2.) return div7_96 * p14 * p1 * q * p2;
            return 0;
        }
    };
    struct KC6 {
        static constexpr f64 Rkern = 2, hfactd = 2.2, norm_3d = 1365. / (512. * pi);
        static f64 f(f64 q) {
            constexpr f64 div1_2 = (1. / 2.), div25_4 = (25. / 4.);
            f64 p1 = (1 - q * div1_2), p2 = (1 + q * 4 + div25_4 * q * q + 4 * q * q * q);
            p1 *= p1;
            p1 *= p1;
            p1 *= p1;
            if (q < 2.) return p1 * p2;
            return 0;
        }
        static f64 df(f64 q) {
            constexpr f64 div11_512 = (11. / 512.);
            f64 p1 = (-2 + q), p2 = 2 + 7 * q + 8 * q * q;
            f64 p12 = p1 * p1, p14 = p12 * p12, p17 = p12 * p14 * p1;
            if (q < 2.) return div11_512 * p17 * q * p2;
            return 0;
        }
    };

    template<class K>
    struct Kern {
        static constexpr f64 Rkern = K::Rkern, hfactd = K::hfactd;
        static f64 W_3d(f64 r, f64 h) { return K::norm_3d * K::f(r / h) / (h * h * h); }
        static f64 dW_3d(f64 r, f64 h) { return K::norm_3d * K::df(r / h) / (h * h * h * h); }
        static f64 dhW_3d(f64 r, f64 h) {
            return -(K::norm_3d) * (3 * K::f(r / h) + (r / h) * K::df(r / h)) / (h * h * h * h);
        }
    };

    // density.hpp:22-41
    inline f64 rho_h(f64 m, f64 h, f64 hfact) { return m * (hfact / h) * (hfact / h) * (hfact / h); }
    inline f64 newton_iterate_new_h(f64 rho_ha, f64 rho_sum, f64 sumdWdh, f64 h_a) {
        f64 f_iter  = rho_sum - rho_ha;
        f64 df_iter = sumdWdh + 3 * rho_ha / h_a;
        return h_a - f_iter / df_iter;
    }

    // shambackends/math.hpp:840-873
    inline f64 inv_sat_positive(f64 v) { return (v >= 1e-9) ? 1. / v : 0.; }
    inline f64 inv_sat_zero(f64 v) { return (v != 0. && v == v) ? 1. / v : 0.; }

    // ------------------------------------------------------------------------------------
    // Morton codes.  shammath/sfc/morton.hpp:28-230, sfc/bmi.hpp:30-51,
    // CoordRangeTransform.cpp:168-184 (fact = delta / f64(val_count), mode multiply),
    // CoordRangeTransform.hpp:115-122 (reverse = convert<int>((c - min) / fact)),
    // AABB.hpp:320-322 (clamp_coord), MortonCodeSet.cpp:62-118.
    // ------------------------------------------------------------------------------------
    template<class M>
    struct MortonTraits;
    template<>
    struct MortonTraits<u32> {
        using icoord                  = std::uint16_t; // u16_3
        static constexpr u32 max_val  = 1024 - 1;
        static constexpr u32 val_cnt  = 1024;
        static constexpr u32 err_code = 4294967295U;
        static constexpr int bits     = 32;
        static u32 expand(u32 x) {
            x &= 0x3ffU;
            x = (x | x << 16U) & 0x30000ffU;
            x = (x | x << 8U) & 0x300f00fU;
            x = (x | x << 4U) & 0x30c30c3U;
            x = (x | x << 2U) & 0x9249249U;
            return x;
        }
    };
    template<>
    struct MortonTraits<u64> {
        using icoord                  = std::uint32_t; // u32_3
        static constexpr u64 max_val  = 2097152 - 1;
        static constexpr u64 val_cnt  = 2097152;
        static constexpr u64 err_code = 18446744073709551615UL;
        static constexpr int bits     = 64;
        static u64 expand(u64 x) {
            x &= 0x1fffffUll;
            x = (x | x << 32Ull) & 0x1f00000000ffffUll;
            x = (x | x << 16Ull) & 0x1f0000ff0000ffUll;
            x = (x | x << 8Ull) & 0x100f00f00f00f00fUll;
            x = (x | x << 4Ull) & 0x10c30c30c30c30c3Ull;
            x = (x | x << 2Ull) & 0x1249249249249249Ull;
            return x;
        }
    };

    template<class M>
    inline M morton_of(vec3 r, vec3 bmin, vec3 bmax) {
        using T = MortonTraits<M>;
        using I = typename T::icoord;
        // clamp_coord: sycl::clamp(coord, lower, upper) = fmin(fmax(c, lo), hi)
        vec3 c{std::fmin(std::fmax(r.x, bmin.x), bmax.x),
               std::fmin(std::fmax(r.y, bmin.y), bmax.y),
               std::fmin(std::fmax(r.z, bmin.z), bmax.z)};
        f64 vc = (f64) T::val_cnt;
        vec3 fact{(bmax.x - bmin.x) / vc, (bmax.y - bmin.y) / vc, (bmax.z - bmin.z) / vc};
        I ix = (I) ((c.x - bmin.x) / fact.x);
        I iy = (I) ((c.y - bmin.y) / fact.y);
        I iz = (I) ((c.z - bmin.z) / fact.z);
        // to_morton_grid clamp (min(max(v,0),max_val)) : morton.hpp:213-224
        I mv = (I) T::max_val;
        ix   = std::min(std::max(ix, (I) 0), mv);
        iy   = std::min(std::max(iy, (I) 0), mv);
        iz   = std::min(std::max(iz, (I) 0), mv);
        M xx = T::expand((M) ix), yy = T::expand((M) iy), zz = T::expand((M) iz);
        return xx * 4 + yy * 2 + zz;
    }

    // BuildTrees.cpp:38-55 : bbox of positions widened by one ulp outward with nextafter
    inline void bbox_nextafter(
        const f64 *x, const f64 *y, const f64 *z, u32 n, vec3 &bmin, vec3 &bmax) {
        f64 inf = std::numeric_limits<f64>::infinity();
        vec3 lo{inf, inf, inf}, hi{-inf, -inf, -inf};
        for (u32 i = 0; i < n; i++) {
            lo = vmin(lo, {x[i], y[i], z[i]});
            hi = vmax(hi, {x[i], y[i], z[i]});
        }
        bmin = {std::nextafter(lo.x, -inf), std::nextafter(lo.y, -inf), std::nextafter(lo.z, -inf)};
        bmax = {std::nextafter(hi.x, inf), std::nextafter(hi.y, inf), std::nextafter(hi.z, inf)};
    }

    // shambase/integer.hpp:92-103
    inline u32 roundup_pow2(u32 v) {
        v--;
        v |= v >> 1;
        v |= v >> 2;
        v |= v >> 4;
        v |= v >> 8;
        v |= v >> 16;
        v++;
        return v;
    }

    // ------------------------------------------------------------------------------------
    // Bitonic sort-by-key replay.  shamalgs/src/details/algorithm/bitonicSort_updated_usm.cpp
    // :26-396 with MaxStencilSize = 16 (sort_by_keys.cpp:35-43).  Within one launch the work
    // items touch disjoint elements, so replaying work items sequentially gives exactly the
    // device result, including the (unstable) order of equal keys.
    // ------------------------------------------------------------------------------------
    template<class K>
    inline void bitonic_orderV(K *x, u32 *vx, u32 a, u32 b, bool reverse) {
        bool swap = reverse ^ (x[a] < x[b]);
        K ka = x[a], kb = x[b];
        u32 va = vx[a], vb = vx[b];
        x[a]  = swap ? kb : ka;
        x[b]  = swap ? ka : kb;
        vx[a] = swap ? vb : va;
        vx[b] = swap ? va : vb;
    }
    template<class K>
    inline void bitonic_stencil(K *x, u32 *vx, u32 a, u32 S, bool reverse) {
        if (S < 2) return;
        u32 h = S / 2;
        for (u32 i = 0; i < h; i++) bitonic_orderV(x, vx, a + i, a + i + h, reverse);
        bitonic_stencil(x, vx, a, h, reverse);
        bitonic_stencil(x, vx, a + h, h, reverse);
    }
    template<class K>
    inline void bitonic_sort_by_key(K *m, u32 *id, u32 len) {
        if (len == 0 || (len & (len - 1)) != 0)
            throw std::invalid_argument("bitonic sort length must be a power of two");
        for (u32 length = 1; length < len; length <<= 1) {
            u32 inc = length;
            while (inc > 0) {
                u32 ninc = (inc >= 8) ? 4 : (inc >= 4) ? 3 : (inc >= 2) ? 2 : 1;
                u32 S    = 1u << ninc;
                u32 _inc = inc >> (ninc - 1);
                u32 _dir = length << 1;
                u32 nthreads = len >> ninc;
                K xs[16];
                u32 is[16];
                for (u32 t = 0; t < nthreads; t++) {
                    i32 low      = (i32) (t & (_inc - 1));
                    i32 i        = (i32) (((t - (u32) low) << ninc) + (u32) low);
                    bool reverse = ((_dir & (u32) i) == 0);
                    for (u32 k = 0; k < S; k++) {
                        xs[k] = m[k * _inc + i];
                        is[k] = id[k * _inc + i];
                    }
                    bitonic_stencil(xs, is, 0, S, reverse);
                    for (u32 k = 0; k < S; k++) {
                        m[k * _inc + i]  = xs[k];
                        id[k * _inc + i] = is[k];
                    }
                }
                inc >>= ninc;
            }
        }
    }

    // ------------------------------------------------------------------------------------
    // Leaf reduction.  shamtree/src/kernels/reduction_alg.cpp:62-85 (split table),
    // :196-244 (iteration, NEW_BEHAVIOR), :274-295 (update_morton_buf), :357-373 (indexmap),
    // :564-585 (remap), MortonReducedSet.cpp:24-73, shambackends/math.hpp:782-828.
    // ------------------------------------------------------------------------------------
    template<class M>
    inline i32 clz_m(M v) {
        constexpr int B = MortonTraits<M>::bits;
        if (v == 0) return B;
        if constexpr (B == 32) return __builtin_clz(v);
        else return __builtin_clzll(v);
    }
    template<class M>
    inline i32 karras_delta(i32 x, i32 y, u32 morton_length, const M *m) {
        // (y > morton_length - 1 || y < 0): y is promoted to u32 in the first comparison
        return ((u32) y > morton_length - 1 || y < 0) ? -1 : clz_m<M>(m[x] ^ m[y]);
    }

    // returns buf_reduc_index_map (leaf_count + 2 entries: starts..., cnt_obj, 0)
    template<class M>
    inline std::vector<u32> reduction_index_map(const M *m, u32 cnt, u32 level) {
        std::vector<u32> s_in(cnt), s_out(cnt);
        for (u32 i = 0; i < cnt; i++) s_in[i] = (i > 0) ? (m[i - 1] != m[i] ? 1 : 0) : 1;
        for (u32 iter = 1; iter <= level; iter++) {
            for (u32 i = 0; i < cnt; i++) {
                u32 before1 = i - 1;
                while (before1 <= cnt - 1 && !s_in[before1]) before1--;
                u32 before2 = before1 - 1;
                while (before2 <= cnt - 1 && !s_in[before2]) before2--;
                u32 next1 = i + 1;
                while (next1 <= cnt - 1 && !s_in[next1]) next1++;
                i32 delt_0  = karras_delta<M>((i32) i, (i32) next1, cnt, m);
                i32 delt_m  = karras_delta<M>((i32) i, (i32) before1, cnt, m);
                i32 delt_mm = karras_delta<M>((i32) before1, (i32) before2, cnt, m);
                s_out[i] = (!(delt_0 < delt_m && delt_mm < delt_m) && s_in[i]) ? 1 : 0;
            }
            std::swap(s_in, s_out);
        }
        std::vector<u32> map;
        for (u32 i = 0; i < cnt; i++)
            if (s_in[i]) map.push_back(i);
        map.push_back(cnt);
        map.push_back(0);
        return map;
    }

    // ------------------------------------------------------------------------------------
    // Karras 2012 radix tree.  shamtree/src/KarrasRadixTree.cpp:45-143
    // ------------------------------------------------------------------------------------
    template<class M>
    inline void karras(
        const M *morton, u32 L, u32 *lchild, u32 *rchild, u8 *lflag, u8 *rflag, u32 *endrange) {
        if (L <= 1) return;
        u32 I = L - 1;
        for (u32 iu = 0; iu < I; iu++) {
            i32 i     = (i32) iu;
            auto D    = [&](i32 a, i32 b) { return karras_delta<M>(a, b, L, morton); };
            int ddelta = D(i, i + 1) - D(i, i - 1);
            int d      = (ddelta == 0) ? 0 : ((ddelta > 0) ? 1 : -1);
            int delta_min = D(i, i - d);
            int lmax      = 2;
            while (D(i, i + lmax * d) > delta_min) lmax *= 2;
            int l = 0, t = lmax / 2;
            while (t > 0) {
                if (D(i, i + (l + t) * d) > delta_min) l = l + t;
                t = t / 2;
            }
            u32 j       = (u32) (i + l * d);
            endrange[iu] = j;
            int delta_node = D(i, (i32) j);
            int s          = 0;
            float div      = 2;
            t              = (int) std::ceil((float) l / div);
            while (true) {
                int tmp_ = i + (s + t) * d;
                if (D(i, tmp_) > delta_node) s = s + t;
                if (t <= 1) break;
                div *= 2;
                t = (int) std::ceil((float) l / div);
            }
            int gamma  = i + s * d + std::min(d, 0);
            lchild[iu] = (u32) gamma;
            lflag[iu]  = (std::min(iu, j) == (u32) gamma) ? 1 : 0;
            rchild[iu] = (u32) (gamma + 1);
            rflag[iu]  = (std::max(iu, j) == (u32) (gamma + 1)) ? 1 : 0;
        }
    }

    // ------------------------------------------------------------------------------------
    // Compressed-leaf BVH.  shamtree/src/CompressedLeafBVH.cpp:31-95,
    // KarrasRadixTreeAABB.cpp:32-140, KarrasRadixTreeField.hpp:133-222,
    // KarrasTreeTraverser.hpp:55-118, CLBVHObjectIterator.hpp:73-170, LeafCellIterator.hpp:40-54
    // ------------------------------------------------------------------------------------
    template<class M>
    struct CLBVH {
        u32 cnt_obj = 0, morton_count = 0;
        vec3 bmin{}, bmax{};
        std::vector<M> codes;          // unsorted, padded to morton_count with err_code
        std::vector<M> sorted_codes;   // sorted (bitonic order)
        std::vector<u32> sort_index;   // map_morton_id_to_obj_id
        std::vector<u32> reduc_map;    // L + 2
        std::vector<M> leaf_codes;     // L
        u32 L = 0, I = 0;
        std::vector<u32> lchild, rchild, endrange;
        std::vector<u8> lflag, rflag;
        std::vector<vec3> aabb_min, aabb_max; // I + L

        u32 offset_leaf() const { return I; }
        u32 left(u32 id) const { return lchild[id] + I * (u32) lflag[id]; }
        u32 right(u32 id) const { return rchild[id] + I * (u32) rflag[id]; }
        bool is_leaf(u32 id) const { return id >= I; }

        void build_no_aabb(const f64 *x, const f64 *y, const f64 *z, u32 n, vec3 lo, vec3 hi,
                           u32 level, u32 forced_morton_count = 0) {
            if (n == 0) throw std::invalid_argument("obj_cnt is 0, cannot build a CompressedLeafBVH");
            cnt_obj      = n;
            // CompressedLeafBVH.cpp:46 pads to roundup_pow2; the unit tests may force a count
            morton_count = forced_morton_count ? forced_morton_count : roundup_pow2(n);
            bmin = lo;
            bmax = hi;
            codes.assign(morton_count, MortonTraits<M>::err_code);
            for (u32 i = 0; i < n; i++) codes[i] = morton_of<M>({x[i], y[i], z[i]}, lo, hi);
            sorted_codes = codes;
            sort_index.resize(morton_count);
            for (u32 i = 0; i < morton_count; i++) sort_index[i] = i;
            bitonic_sort_by_key(sorted_codes.data(), sort_index.data(), morton_count);
            reduc_map = reduction_index_map<M>(sorted_codes.data(), cnt_obj, level);
            L         = (u32) reduc_map.size() - 2;
            if (L == 0) throw std::runtime_error("0 leaf tree cannot exists");
            leaf_codes.resize(L);
            for (u32 i = 0; i < L; i++) leaf_codes[i] = sorted_codes[reduc_map[i]];
            I = L - 1;
            lchild.assign(I, 0);
            rchild.assign(I, 0);
            endrange.assign(I, 0);
            lflag.assign(I, 0);
            rflag.assign(I, 0);
            karras<M>(leaf_codes.data(), L, lchild.data(), rchild.data(), lflag.data(), rflag.data(),
                      endrange.data());
        }

        template<class F>
        void for_each_in_leaf(u32 leaf, F &&f) const {
            for (u32 s = reduc_map[leaf]; s < reduc_map[leaf + 1]; s++) f(sort_index[s]);
        }

        // tree_depth = bitsizeof(Tmorton) passes of the child union (KarrasRadixTree.cpp:147-150)
        template<class T, class Comb>
        void propagate_up(std::vector<T> &field, Comb comb) const {
            if (I == 0) return;
            for (int pass = 0; pass < MortonTraits<M>::bits; pass++)
                for (u32 g = 0; g < I; g++) field[g] = comb(field[left(g)], field[right(g)]);
        }

        void compute_aabb(const f64 *x, const f64 *y, const f64 *z) {
            f64 mx = std::numeric_limits<f64>::max();
            aabb_min.assign(I + L, vec3{0, 0, 0});
            aabb_max.assign(I + L, vec3{0, 0, 0});
            for (u32 l = 0; l < L; l++) {
                vec3 lo{mx, mx, mx}, hi{-mx, -mx, -mx};
                for_each_in_leaf(l, [&](u32 o) {
                    lo = vmin(lo, {x[o], y[o], z[o]});
                    hi = vmax(hi, {x[o], y[o], z[o]});
                });
                aabb_min[I + l] = lo;
                aabb_max[I + l] = hi;
            }
            propagate_up(aabb_min, [](vec3 a, vec3 b) { return vmin(a, b); });
            propagate_up(aabb_max, [](vec3 a, vec3 b) { return vmax(a, b); });
        }

        // compute_tree_field_max_field: KarrasRadixTreeField.hpp:188-222
        std::vector<f64> field_max(const f64 *field) const {
            std::vector<f64> out(I + L, 0.);
            for (u32 l = 0; l < L; l++) {
                f64 v = std::numeric_limits<f64>::lowest();
                for_each_in_leaf(l, [&](u32 o) { v = std::fmax(v, field[o]); });
                out[I + l] = v;
            }
            propagate_up(out, [](f64 a, f64 b) { return std::fmax(a, b); });
            return out;
        }

        // KarrasTreeTraverser.hpp:69-118 stack traversal (left child visited first)
        template<class Crit, class Found>
        void traverse(Crit &&crit, Found &&found) const {
            constexpr u32 depth = 64 + 1;
            u32 stack[depth];
            u32 cur         = depth - 1;
            stack[cur]      = 0;
            while (cur < depth) {
                u32 node = stack[cur];
                cur++;
                if (crit(node, aabb_min[node], aabb_max[node])) {
                    if (is_leaf(node)) {
                        found(node);
                    } else {
                        stack[cur - 1] = right(node);
                        cur--;
                        stack[cur - 1] = left(node);
                        cur--;
                    }
                }
            }
        }
    };

    // geometry_utils.hpp:74-81, :125-135
    inline bool in_range_incl_max(vec3 p, vec3 lo, vec3 hi) {
        return (lo.x <= p.x) && (p.x <= hi.x) && (lo.y <= p.y) && (p.y <= hi.y) && (lo.z <= p.z) &&
               (p.z <= hi.z);
    }
    inline bool cella_neigh_b(vec3 amin, vec3 amax, vec3 bmin, vec3 bmax) {
        return (std::fmax(amin.x, bmin.x) <= std::fmin(amax.x, bmax.x)) &&
               (std::fmax(amin.y, bmin.y) <= std::fmin(amax.y, bmax.y)) &&
               (std::fmax(amin.z, bmin.z) <= std::fmin(amax.z, bmax.z));
    }

    // TreeTraversal.hpp:375-485 ObjectCache (CSR)
    struct ObjectCache {
        std::vector<u32> cnt, offset, ids;
    };
    inline ObjectCache make_cache(std::vector<u32> &&cnt) {
        ObjectCache c;
        c.cnt = std::move(cnt);
        c.offset.resize(c.cnt.size());
        u32 acc = 0;
        for (size_t i = 0; i < c.cnt.size(); i++) {
            c.offset[i] = acc;
            acc += c.cnt[i];
        }
        c.ids.resize(acc);
        return c;
    }

    // ------------------------------------------------------------------------------------
    // Two-stage neighbour cache.  shammodels/sph/src/modules/NeighbourCache.cpp:222-604
    //   rint_tree = tree max-h field already multiplied by htol (Solver.cpp:1322-1356)
    //   obj_cnt = local particle count (first obj_cnt entries of the merged arrays)
    // ------------------------------------------------------------------------------------
    struct NeighOut {
        ObjectCache leaf_cache;     // per leaf: list of neighbour leaf node ids (I+leaf)
        std::vector<u32> owner;     // per local particle: owner leaf index
        ObjectCache part_cache;     // per local particle: neighbour ids in merged arrays
    };

    template<class M>
    inline NeighOut neigh_cache_2stages(const CLBVH<M> &t, const std::vector<f64> &rint_tree,
                                        const f64 *x, const f64 *y, const f64 *z, const f64 *h,
                                        u32 obj_cnt, f64 h_tolerance, f64 Rkern) {
        NeighOut o;
        u32 off = t.offset_leaf();
        auto leaf_walk = [&](u32 gid, auto &&on_leaf) {
            f64 leaf_a_rint = rint_tree[off + gid] * Rkern;
            vec3 amin = t.aabb_min[off + gid], amax = t.aabb_max[off + gid];
            vec3 amin_ext{amin.x - leaf_a_rint, amin.y - leaf_a_rint, amin.z - leaf_a_rint};
            vec3 amax_ext{amax.x + leaf_a_rint, amax.y + leaf_a_rint, amax.z + leaf_a_rint};
            t.traverse(
                [&](u32 node, vec3 nlo, vec3 nhi) {
                    f64 r = rint_tree[node] * Rkern;
                    vec3 elo{nlo.x - r, nlo.y - r, nlo.z - r};
                    vec3 ehi{nhi.x + r, nhi.y + r, nhi.z + r};
                    return cella_neigh_b(amin, amax, elo, ehi) ||
                           cella_neigh_b(amin_ext, amax_ext, nlo, nhi);
                },
                on_leaf);
        };
        {
            std::vector<u32> cnt(t.L, 0);
#pragma omp parallel for schedule(dynamic, 64)
            for (u32 g = 0; g < t.L; g++) {
                u32 c = 0;
                leaf_walk(g, [&](u32) { c++; });
                cnt[g] = c;
            }
            o.leaf_cache = make_cache(std::move(cnt));
#pragma omp parallel for schedule(dynamic, 64)
            for (u32 g = 0; g < t.L; g++) {
                u32 c = o.leaf_cache.offset[g];
                leaf_walk(g, [&](u32 lb) { o.leaf_cache.ids[c++] = lb; });
            }
        }
        o.owner.assign(obj_cnt, 0);
        bool not_found = false;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < obj_cnt; a++) {
            vec3 ra{x[a], y[a], z[a]};
            u32 found = (u32) std::numeric_limits<i32>::max();
            t.traverse([&](u32, vec3 lo, vec3 hi) { return in_range_incl_max(ra, lo, hi); },
                       [&](u32 lb) { found = lb - off; });
            if (found >= t.L) {
#pragma omp critical
                not_found = true;
            }
            o.owner[a] = found;
        }
        if (not_found) throw std::runtime_error("neighbour cache: particle outside every leaf");
        const f64 Rker2 = Rkern * Rkern;
        auto part_walk = [&](u32 a, auto &&on_b) {
            f64 rint_a = h[a] * h_tolerance;
            vec3 ra{x[a], y[a], z[a]};
            u32 lo = o.owner[a];
            u32 s = o.leaf_cache.offset[lo], e = s + o.leaf_cache.cnt[lo];
            for (u32 k = s; k < e; k++) {
                u32 lb = o.leaf_cache.ids[k];
                t.for_each_in_leaf(lb - off, [&](u32 b) {
                    vec3 dr    = ra - vec3{x[b], y[b], z[b]};
                    f64 rab2   = dot(dr, dr);
                    f64 rint_b = h[b] * h_tolerance;
                    bool no_interact = rab2 > rint_a * rint_a * Rker2 && rab2 > rint_b * rint_b * Rker2;
                    if (!no_interact) on_b(b);
                });
            }
        };
        std::vector<u32> pc(obj_cnt, 0);
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < obj_cnt; a++) {
            u32 c = 0;
            part_walk(a, [&](u32) { c++; });
            pc[a] = c;
        }
        o.part_cache = make_cache(std::move(pc));
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < obj_cnt; a++) {
            u32 c = o.part_cache.offset[a];
            part_walk(a, [&](u32 b) { o.part_cache.ids[c++] = b; });
        }
        return o;
    }

    // ------------------------------------------------------------------------------------
    // Smoothing-length Newton step.  IterateSmoothingLengthDensity.cpp:28-120
    // ------------------------------------------------------------------------------------
    template<class K>
    inline void iterate_h(u32 n, const ObjectCache &c, const f64 *x, const f64 *y, const f64 *z,
                          const f64 *h_old, f64 *h_new, f64 *eps, f64 gpart_mass, f64 h_evol_max,
                          f64 h_evol_iter_max) {
        using KK = Kern<K>;
        const f64 Rkern = K::Rkern;
        f64 h_max_evol_p = h_evol_iter_max, h_max_evol_m = 1 / h_evol_iter_max;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            if (!(eps[a] > 1e-6)) continue;
            vec3 ra{x[a], y[a], z[a]};
            f64 h_a = h_new[a];
            f64 dint = h_a * h_a * Rkern * Rkern;
            f64 rho_sum = 0, sumdWdh = 0;
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b    = c.ids[k];
                vec3 dr  = ra - vec3{x[b], y[b], z[b]};
                f64 rab2 = dot(dr, dr);
                if (rab2 > dint) continue;
                f64 rab = std::sqrt(rab2);
                rho_sum += gpart_mass * KK::W_3d(rab, h_a);
                sumdWdh += gpart_mass * KK::dhW_3d(rab, h_a);
            }
            f64 rho_ha = rho_h(gpart_mass, h_a, K::hfactd);
            f64 new_h  = newton_iterate_new_h(rho_ha, rho_sum, sumdWdh, h_a);
            if (new_h < h_a * h_max_evol_m) new_h = h_max_evol_m * h_a;
            if (new_h > h_a * h_max_evol_p) new_h = h_max_evol_p * h_a;
            f64 ha_0 = h_old[a];
            if (new_h < ha_0 * h_evol_max) {
                h_new[a] = new_h;
                eps[a]   = std::fabs(new_h - h_a) / ha_0;
            } else {
                h_new[a] = ha_0 * h_evol_max;
                eps[a]   = -1;
            }
        }
    }

    // NodeComputeOmega: ComputeOmega.cpp:25-74
    template<class K>
    inline void compute_omega(u32 n, const ObjectCache &c, const f64 *x, const f64 *y,
                              const f64 *z, const f64 *h, f64 *omega, f64 pmass) {
        using KK = Kern<K>;
        const f64 Rkern = K::Rkern;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            vec3 ra{x[a], y[a], z[a]};
            f64 h_a = h[a];
            f64 dint = h_a * h_a * Rkern * Rkern;
            f64 rho_sum = 0, part_omega_sum = 0;
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b    = c.ids[k];
                vec3 dr  = ra - vec3{x[b], y[b], z[b]};
                f64 rab2 = dot(dr, dr);
                if (rab2 > dint) continue;
                f64 rab = std::sqrt(rab2);
                rho_sum += pmass * KK::W_3d(rab, h_a);
                part_omega_sum += pmass * KK::dhW_3d(rab, h_a);
            }
            f64 rho_ha = rho_h(pmass, h_a, K::hfactd);
            omega[a]   = 1 + (h_a / (3 * rho_ha)) * part_omega_sum;
        }
    }

    // ComputeEOSAdiabatic (ComputeEos.cpp:152-247; shamphys/eos.hpp:44-54), isothermal
    // (eos.hpp:31-35; ComputeEos.cpp:656-665), locally isothermal LP07 (eos.hpp:86-98;
    // ComputeEos.cpp:726-798)
    enum EosKind : int { EOS_ADIABATIC = 0, EOS_ISOTHERMAL = 1, EOS_LP07 = 2 };
    struct EosParams {
        int kind   = EOS_ADIABATIC;
        f64 gamma  = 5. / 3.;
        f64 cs     = 1.;   // isothermal
        f64 cs0    = 1.;   // LP07
        f64 q      = 0.5;  // LP07
        f64 r0     = 1.;   // LP07
    };
    template<class K>
    inline void compute_eos(u32 nm, const f64 *h, const f64 *u, const f64 *x, const f64 *y,
                            const f64 *z, f64 *P, f64 *cs, f64 pmass, const EosParams &e) {
#pragma omp parallel for
        for (u32 i = 0; i < nm; i++) {
            f64 rho = rho_h(pmass, h[i], K::hfactd);
            if (e.kind == EOS_ADIABATIC) {
                f64 P_a = (e.gamma - 1) * rho * u[i];
                P[i]    = P_a;
                cs[i]   = std::sqrt(e.gamma * P_a / rho);
            } else if (e.kind == EOS_ISOTHERMAL) {
                P[i]  = e.cs * e.cs * rho;
                cs[i] = e.cs;
            } else {
                // LP07 (ComputeEos.cpp:726-798): cs^2 = cs0^2 * pow(|r|^2 / r0^2, -q)
                f64 r0sq  = e.r0 * e.r0;
                f64 mq    = -e.q;
                f64 Rsq   = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
                f64 cs_sq = (e.cs0 * e.cs0) * std::pow(Rsq / r0sq, mq);
                cs[i]     = std::sqrt(cs_sq);
                P[i]      = cs_sq * rho;
            }
        }
    }

    // ------------------------------------------------------------------------------------
    // Force + du/dt, constant AV.  UpdateDerivs.cpp:88-288; math/forces.hpp:34-226;
    // math/q_ab.hpp:36-39
    // ------------------------------------------------------------------------------------
    struct AVParams {
        f64 alpha_u = 1.0, alpha_AV = 1.0, beta_AV = 2.0;
    };
    template<class K>
    inline void update_derivs_constant_av(u32 n, const ObjectCache &c, const f64 *x,
                                          const f64 *y, const f64 *z, const f64 *vx,
                                          const f64 *vy, const f64 *vz, const f64 *h,
                                          const f64 *omega, const f64 *u, const f64 *P,
                                          const f64 *cs, f64 *ax, f64 *ay, f64 *az, f64 *du,
                                          f64 pmass, AVParams av,
                                          const f64 *alpha = nullptr) {
        using KK = Kern<K>;
        constexpr f64 Rker2 = K::Rkern * K::Rkern;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            f64 h_a = h[a];
            vec3 xyz_a{x[a], y[a], z[a]}, v_a{vx[a], vy[a], vz[a]};
            f64 P_a = P[a], omega_a = omega[a];
            const f64 u_a = u[a];
            f64 rho_a = rho_h(pmass, h_a, K::hfactd);
            f64 rho_a_sq = rho_a * rho_a;
            f64 rho_a_inv = 1. / rho_a;
            f64 omega_a_rho_a_inv = 1 / (omega_a * rho_a);
            f64 cs_a = cs[a];
            vec3 force{0, 0, 0};
            f64 tdu = 0;
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b    = c.ids[k];
                vec3 dr  = xyz_a - vec3{x[b], y[b], z[b]};
                f64 rab2 = dot(dr, dr);
                f64 h_b  = h[b];
                if (rab2 > h_a * h_a * Rker2 && rab2 > h_b * h_b * Rker2) continue;
                f64 rab = std::sqrt(rab2);
                vec3 v_b{vx[b], vy[b], vz[b]};
                const f64 u_b = u[b];
                f64 rho_b = rho_h(pmass, h_b, K::hfactd);
                f64 P_b = P[b], omega_b = omega[b], cs_b = cs[b];
                f64 Fab_a = KK::dW_3d(rab, h_a), Fab_b = KK::dW_3d(rab, h_b);
                vec3 v_ab      = v_a - v_b;
                vec3 r_ab_unit = dr * inv_sat_positive(rab);
                f64 v_ab_r_ab  = dot(v_ab, r_ab_unit);
                f64 abs_v      = std::fabs(v_ab_r_ab);
                // per-particle alpha: NodeUpdateDerivsVaryingAlphaAV.cpp:60-150
                f64 alpha_a = alpha ? alpha[a] : av.alpha_AV;
                f64 alpha_b = alpha ? alpha[b] : av.alpha_AV;
                f64 vsig_a  = alpha_a * cs_a + av.beta_AV * abs_v;
                f64 vsig_b  = alpha_b * cs_b + av.beta_AV * abs_v;
                f64 rho_avg = (rho_a + rho_b) * 0.5;
                f64 vsig_u  = std::sqrt(std::fabs(P_a - P_b) / rho_avg);
                f64 qa_ab = std::fmax(-0.5 * rho_a * vsig_a * v_ab_r_ab, 0.);
                f64 qb_ab = std::fmax(-0.5 * rho_b * vsig_b * v_ab_r_ab, 0.);
                // add_to_derivs_sph_artif_visco_cond (forces.hpp:172-226)
                f64 AV_P_a = P_a + qa_ab, AV_P_b = P_b + qb_ab;
                vec3 nab_a = r_ab_unit * Fab_a, nab_b = r_ab_unit * Fab_b;
                f64 sub_a = rho_a_sq * omega_a, sub_b = (rho_b * rho_b) * omega_b;
                vec3 acc_a = (AV_P_a * inv_sat_zero(sub_a)) * nab_a;
                vec3 acc_b = (AV_P_b * inv_sat_zero(sub_b)) * nab_b;
                force = force + (-pmass) * (acc_a + acc_b);
                tdu += AV_P_a * (omega_a_rho_a_inv * rho_a_inv) * pmass * dot(v_ab, nab_a);
                tdu += pmass * av.alpha_u * vsig_u * (u_a - u_b) * 0.5 *
                       (Fab_a * omega_a_rho_a_inv + Fab_b / (rho_b * omega_b));
            }
            ax[a] = force.x;
            ay[a] = force.y;
            az[a] = force.z;
            du[a] = tdu;
        }
    }

    // vsig for the CFL: Solver.cpp:2716-2789 (alpha = 1, beta = 2 hard-coded there)
    template<class K>
    inline void compute_vsig(u32 n, const ObjectCache &c, const f64 *x, const f64 *y,
                             const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz,
                             const f64 *h, const f64 *cs, f64 *vsig) {
        constexpr f64 Rker2 = K::Rkern * K::Rkern;
        const f64 alpha_AV = 1.0, beta_AV = 2.0;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            f64 h_a = h[a];
            vec3 xyz_a{x[a], y[a], z[a]}, v_a{vx[a], vy[a], vz[a]};
            f64 cs_a = cs[a];
            f64 vsig_max = 0;
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b    = c.ids[k];
                vec3 dr  = xyz_a - vec3{x[b], y[b], z[b]};
                f64 rab2 = dot(dr, dr);
                f64 h_b  = h[b];
                if (rab2 > h_a * h_a * Rker2 && rab2 > h_b * h_b * Rker2) continue;
                f64 rab        = std::sqrt(rab2);
                vec3 v_ab      = v_a - vec3{vx[b], vy[b], vz[b]};
                vec3 r_ab_unit = dr / rab;
                if (rab < 1e-9) r_ab_unit = {0, 0, 0};
                f64 v_ab_r_ab = dot(v_ab, r_ab_unit);
                f64 vsig_a    = alpha_AV * cs_a + beta_AV * std::fabs(v_ab_r_ab);
                vsig_max      = std::fmax(vsig_max, vsig_a);
            }
            vsig[a] = vsig_max;
        }
    }

    // ------------------------------------------------------------------------------------
    // Varying artificial viscosity, Cullen & Dehnen (2010) (AVConfig VaryingCD10)
    // ------------------------------------------------------------------------------------
    inline vec3 cross(vec3 a, vec3 b) {
        return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    }

    // div v and curl v: DiffOperator.cpp:24-145 (update_divv) and :147-268 (update_curlv),
    // same loop and criterion; the two sums are independent, so one loop gives both
    template<class K>
    inline void compute_divv_curlv(u32 n, const ObjectCache &c, const f64 *x, const f64 *y,
                                   const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz,
                                   const f64 *h, const f64 *omega, f64 pmass, f64 *divv,
                                   f64 *curlx, f64 *curly, f64 *curlz) {
        using KK = Kern<K>;
        constexpr f64 Rker2 = K::Rkern * K::Rkern;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            f64 h_a = h[a];
            vec3 xyz_a{x[a], y[a], z[a]}, v_a{vx[a], vy[a], vz[a]};
            f64 omega_a = omega[a];
            f64 rho_a   = rho_h(pmass, h_a, K::hfactd);
            f64 inv_rho_omega_a = 1. / (omega_a * rho_a);
            f64 sum_nabla_v = 0;
            vec3 sum_cross{0, 0, 0};
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b    = c.ids[k];
                vec3 dr  = xyz_a - vec3{x[b], y[b], z[b]};
                f64 rab2 = dot(dr, dr);
                f64 h_b  = h[b];
                if (rab2 > h_a * h_a * Rker2 && rab2 > h_b * h_b * Rker2) continue;
                f64 rab        = std::sqrt(rab2);
                vec3 v_ab      = v_a - vec3{vx[b], vy[b], vz[b]};
                vec3 r_ab_unit = dr / rab;
                if (rab < 1e-9) r_ab_unit = {0, 0, 0};
                vec3 dWab_a = KK::dW_3d(rab, h_a) * r_ab_unit;
                sum_nabla_v += pmass * dot(v_ab, dWab_a);
                sum_cross = sum_cross + pmass * cross(v_ab, dWab_a);
            }
            divv[a]  = -inv_rho_omega_a * sum_nabla_v;
            curlx[a] = -inv_rho_omega_a * sum_cross.x;
            curly[a] = -inv_rho_omega_a * sum_cross.y;
            curlz[a] = -inv_rho_omega_a * sum_cross.z;
        }
    }

    // 3x3 inverse and product of row-stored matrices (matrix_legacy.hpp:25-95)
    struct mat33 {
        vec3 r[3];
    };
    inline mat33 inv_33(const mat33 &m) {
        f64 a00 = m.r[0].x, a10 = m.r[1].x, a20 = m.r[2].x;
        f64 a01 = m.r[0].y, a11 = m.r[1].y, a21 = m.r[2].y;
        f64 a02 = m.r[0].z, a12 = m.r[1].z, a22 = m.r[2].z;
        f64 det = (-a02 * a11 * a20 + a01 * a12 * a20 + a02 * a10 * a21 - a00 * a12 * a21 -
                   a01 * a10 * a22 + a00 * a11 * a22);
        return {{vec3{-a12 * a21 + a11 * a22, a02 * a21 - a01 * a22, -a02 * a11 + a01 * a12} / det,
                 vec3{a12 * a20 - a10 * a22, -a02 * a20 + a00 * a22, a02 * a10 - a00 * a12} / det,
                 vec3{-a11 * a20 + a10 * a21, a01 * a20 - a00 * a21, -a01 * a10 + a00 * a11} /
                     det}};
    }
    inline mat33 prod_33(const mat33 &A, const mat33 &B) {
        f64 a00 = A.r[0].x, a10 = A.r[1].x, a20 = A.r[2].x;
        f64 a01 = A.r[0].y, a11 = A.r[1].y, a21 = A.r[2].y;
        f64 a02 = A.r[0].z, a12 = A.r[1].z, a22 = A.r[2].z;
        f64 b00 = B.r[0].x, b10 = B.r[1].x, b20 = B.r[2].x;
        f64 b01 = B.r[0].y, b11 = B.r[1].y, b21 = B.r[2].y;
        f64 b02 = B.r[0].z, b12 = B.r[1].z, b22 = B.r[2].z;
        return {{vec3{a00 * b00 + a01 * b10 + a02 * b20, a00 * b01 + a01 * b11 + a02 * b21,
                      a00 * b02 + a01 * b12 + a02 * b22},
                 vec3{a10 * b00 + a11 * b10 + a12 * b20, a10 * b01 + a11 * b11 + a12 * b21,
                      a10 * b02 + a11 * b12 + a12 * b22},
                 vec3{a20 * b00 + a21 * b10 + a22 * b20, a20 * b01 + a21 * b11 + a22 * b21,
                      a20 * b02 + a21 * b12 + a22 * b22}}};
    }

    // d(div v)/dt = div a - (grad v):(grad v)^T from the matrix-corrected gradient estimates
    // (DiffOperatorDtDivv.cpp:89-175, update_dtdivv without div / curl)
    template<class K>
    inline void compute_dtdivv(u32 n, const ObjectCache &c, const f64 *x, const f64 *y,
                               const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz,
                               const f64 *ax, const f64 *ay, const f64 *az, const f64 *h,
                               f64 pmass, f64 *dtdivv) {
        using KK = Kern<K>;
        constexpr f64 Rker2 = K::Rkern * K::Rkern;
#pragma omp parallel for schedule(dynamic, 256)
        for (u32 a = 0; a < n; a++) {
            f64 h_a = h[a];
            vec3 xyz_a{x[a], y[a], z[a]}, v_a{vx[a], vy[a], vz[a]}, a_a{ax[a], ay[a], az[a]};
            mat33 R{{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}, Rv = R, Ra = R;
            u32 s = c.offset[a], e = s + c.cnt[a];
            for (u32 k = s; k < e; k++) {
                u32 b     = c.ids[k];
                vec3 r_ab = xyz_a - vec3{x[b], y[b], z[b]};
                f64 rab2  = dot(r_ab, r_ab);
                f64 h_b   = h[b];
                if (rab2 > h_a * h_a * Rker2 && rab2 > h_b * h_b * Rker2) continue;
                f64 rab        = std::sqrt(rab2);
                vec3 v_ab      = v_a - vec3{vx[b], vy[b], vz[b]};
                vec3 a_ab      = a_a - vec3{ax[b], ay[b], az[b]};
                vec3 r_ab_unit = r_ab / rab;
                if (rab < 1e-9) r_ab_unit = {0, 0, 0};
                vec3 dWab_a  = KK::dW_3d(rab, h_a) * r_ab_unit;
                vec3 mdWab_b = dWab_a * pmass;
                R.r[0]  = R.r[0] - r_ab.x * mdWab_b;
                R.r[1]  = R.r[1] - r_ab.y * mdWab_b;
                R.r[2]  = R.r[2] - r_ab.z * mdWab_b;
                Rv.r[0] = Rv.r[0] - v_ab * mdWab_b.x;
                Rv.r[1] = Rv.r[1] - v_ab * mdWab_b.y;
                Rv.r[2] = Rv.r[2] - v_ab * mdWab_b.z;
                Ra.r[0] = Ra.r[0] - a_ab * mdWab_b.x;
                Ra.r[1] = Ra.r[1] - a_ab * mdWab_b.y;
                Ra.r[2] = Ra.r[2] - a_ab * mdWab_b.z;
            }
            mat33 inv = inv_33(R);
            mat33 dv  = prod_33(inv, Rv);
            mat33 da  = prod_33(inv, Ra);
            f64 div_a = da.r[0].x + da.r[1].y + da.r[2].z;
            f64 tens  = dv.r[0].x * dv.r[0].x + dv.r[1].x * dv.r[0].y + dv.r[2].x * dv.r[0].z +
                       dv.r[0].y * dv.r[1].x + dv.r[1].y * dv.r[1].y + dv.r[2].y * dv.r[1].z +
                       dv.r[0].z * dv.r[2].x + dv.r[1].z * dv.r[2].y + dv.r[2].z * dv.r[2].z;
            dtdivv[a] = div_a - tens;
        }
    }

    // alpha update: UpdateViscosity.cpp:120-219 (Cullen & Dehnen 2010: limiter from div v and
    // curl v, source from -d(div v)/dt, implicit Euler decay on the time scale h / (c_s sigma));
    // cs is the sound speed saved at the end of the previous step (Solver.cpp:3126-3160), 0
    // before the first step
    struct CD10Params {
        f64 alpha_min, alpha_max, sigma_decay;
    };
    inline void update_alpha_cd10(u32 n, f64 dt, const f64 *cs, const f64 *h, const f64 *alpha,
                                  const f64 *divv, const f64 *curlx, const f64 *curly,
                                  const f64 *curlz, const f64 *dtdivv, CD10Params p,
                                  f64 *alpha_new) {
        const f64 eps = std::numeric_limits<f64>::epsilon();
        for (u32 i = 0; i < n; i++) {
            f64 cs_a = cs[i], h_a = h[i], alpha_a = alpha[i], divv_a = divv[i];
            vec3 curlv_a{curlx[i], curly[i], curlz[i]};
            f64 dtdivv_a        = dtdivv[i];
            f64 vsig            = cs_a;
            f64 inv_tau_a       = vsig * p.sigma_decay / h_a;
            f64 fact_t          = dt * inv_tau_a;
            f64 euler_impl_fact = 1 / (1 + fact_t);
            f64 fac = std::max(-divv_a, 0.);
            fac *= fac;
            f64 traceS  = dot(curlv_a, curlv_a);
            f64 balsara = (fac + traceS > eps) ? fac / (fac + traceS) : 1.;
            f64 A_a     = balsara * std::max(-dtdivv_a, 0.);
            f64 temp    = cs_a * cs_a;
            f64 alpha_loc_a =
                std::min((cs_a > 0) ? 10 * h_a * h_a * A_a / (temp) : p.alpha_min, p.alpha_max);
            alpha_loc_a   = (temp > eps) ? alpha_loc_a : p.alpha_min;
            f64 new_alpha = (alpha_a + alpha_loc_a * fact_t) * euler_impl_fact;
            if (alpha_loc_a > alpha_a) new_alpha = alpha_loc_a;
            alpha_new[i] = new_alpha;
        }
    }

    // ------------------------------------------------------------------------------------
    // Periodic ghosts for one patch.  BasicSPHGhosts.cpp:260-360 (find_interfaces, periodic),
    // :511-579 (gen_id_table_interfaces: is_in_patch_converted, Patch.hpp:210-215 half-open),
    // BasicSPHGhosts.hpp:290-320 (positions + offset), :452-505 (merge: local then ghosts),
    // SPHUtilities.hpp:79-104 (int range = max h * htol * Rkern).
    // Returns the source index of every ghost, ghost offsets written to off_{x,y,z}.
    // ------------------------------------------------------------------------------------
    inline std::vector<u32> periodic_ghost_ids(const f64 *x, const f64 *y, const f64 *z, u32 n,
                                               vec3 box_min, vec3 box_max, f64 int_range,
                                               std::vector<vec3> &offsets) {
        vec3 bsize = box_max - box_min;
        // patch range in object coordinates: u64 patch coords [0, 2^21) mapped by
        // CoordRangeTransform (fact = bsize / 2^21, exact power-of-two scaling)
        vec3 pmin = box_min, pmax = bsize + box_min;
        std::vector<u32> ids;
        offsets.clear();
        for (i32 xo = -1; xo <= 1; xo++)
            for (i32 yo = -1; yo <= 1; yo++)
                for (i32 zo = -1; zo <= 1; zo++) {
                    if (xo == 0 && yo == 0 && zo == 0) continue;
                    vec3 off{xo * bsize.x, yo * bsize.y, zo * bsize.z};
                    // existence test with the patch-tree field (one patch: same int_range)
                    vec3 rlo{pmin.x - int_range, pmin.y - int_range, pmin.z - int_range};
                    vec3 rhi{pmax.x + int_range, pmax.y + int_range, pmax.z + int_range};
                    vec3 slo = pmin + off, shi = pmax + off;
                    vec3 ilo = vmax(rlo, slo), ihi = vmin(rhi, shi);
                    if (!(ihi.x >= ilo.x && ihi.y >= ilo.y && ihi.z >= ilo.z)) continue;
                    // interf_volume = sender ∩ (receiver expanded - offset)
                    vec3 clo = vmax(pmin, rlo - off), chi = vmin(pmax, rhi - off);
                    size_t before = ids.size();
                    for (u32 i = 0; i < n; i++) {
                        bool in = (clo.x <= x[i]) && (x[i] < chi.x) && (clo.y <= y[i]) &&
                                  (y[i] < chi.y) && (clo.z <= z[i]) && (z[i] < chi.z);
                        if (in) {
                            ids.push_back(i);
                            offsets.push_back(off);
                        }
                    }
                    (void) before;
                }
        return ids;
    }

    // sycl_position_modulo: shamrock/src/math/integrators.cpp:206-236
    inline f64 position_modulo(f64 v, f64 bmin, f64 delt) {
        f64 r = v - bmin;
        r     = std::fmod(r, delt);
        r += delt;
        r = std::fmod(r, delt);
        r += bmin;
        return r;
    }

} // namespace orc
