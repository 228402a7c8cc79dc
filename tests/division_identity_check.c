/* Checks x / y == RN(q + RN(x - q y) r) with r = RN(1 / y), q = RN(x r), the FMA form of the
 * division the HIP pair loops use (div_rcp in shamrock_amd/csrc/sph.hip), on random and
 * adversarial operands (significands of all ones, near powers of two, exponents in
 * [-255, 255]).  Prints the number of mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st[2] = {0x12345678abcdefULL, 0x9e3779b97f4a7c15ULL};
static uint64_t next(void) {
    uint64_t s1 = st[0];
    const uint64_t s0 = st[1];
    st[0] = s0;
    s1 ^= s1 << 23;
    st[1] = s1 ^ s0 ^ (s1 >> 17) ^ (s0 >> 26);
    return st[1] + s0;
}
static double make(int e, uint64_t m, int neg) {
    uint64_t b = ((uint64_t) (e + 1023) << 52) | (m & ((1ULL << 52) - 1));
    double d;
    memcpy(&d, &b, 8);
    return neg ? -d : d;
}

int main(int argc, char **argv) {
    long n = argc > 1 ? atol(argv[1]) : 10000000L, bad = 0;
    for (long i = 0; i < n; i++) {
        uint64_t ra = next(), rb = next();
        int ea = (int) ((ra >> 52) & 511) - 255, eb = (int) ((rb >> 52) & 511) - 255;
        uint64_t ma = ra, mb = rb;
        switch (i & 7) {
        case 1: mb = ~0ULL; break;                  /* divisor significand all ones */
        case 2: mb = ~0ULL - (rb & 0xff); break;    /* just below a power of two */
        case 3: mb = rb & 0xff; break;              /* just above a power of two */
        case 4: ma = ~0ULL; break;
        case 5: ma = rb; break;                     /* x and y with the same significand */
        default: break;
        }
        double x = make(ea, ma, (int) (ra >> 63)), y = make(eb, mb, (int) ((rb >> 62) & 1));
        double r = 1.0 / y, q = x * r;
        double e = fma(-q, y, x);
        double d = fma(e, r, q);
        if (d != x / y) {
            if (bad < 5) printf("mismatch x=%a y=%a got=%a want=%a\n", x, y, d, x / y);
            bad++;
        }
    }
    printf("%ld\n", bad);
    return 0;
}
