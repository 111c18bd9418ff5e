/* Checks the lookups' division by the leaf extent through the reciprocal:
 * for an extent e with y = RN(1/e), q = RN(a*y), r = fma(-e, q, a),
 * q' = fma(r, y, q) must equal RN(a/e) for every float a with |a| >= 2^-100
 * (smaller numerators take the division in the kernels).
 *   markstein_check all <e>          every float numerator
 *   markstein_check random <n> <seed> n random numerators for each of 64
 *                                    random extents in [2^-20, 2^10]
 *   markstein_check tiny <e>         every 97th nonzero numerator below
 *                                    2^-100 (which the kernels divide): how
 *                                    many the correction would get wrong
 * Prints the number of mismatches; exit status 1 if any (for tiny: 0). */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static int differs(float a, float e, float y) {
    const float d = a / e;
    if (isinf(d)) return 0;
    const float q = a * y;
    const float r = fmaf(-e, q, a);
    const float m = fmaf(r, y, q);
    return u_of(m) != u_of(d);
}

static int bad(float a, float e, float y) {
    if (!(fabsf(a) >= 0x1p-100f) || isinf(a)) return 0;   /* NaN, tiny, zero, inf: division */
    return differs(a, e, y);
}

typedef struct { float e; uint64_t lo, hi, nbad; } job;

static void *run_all(void *p) {
    job *J = (job *)p;
    const float y = 1.0f / J->e;
    for (uint64_t u = J->lo; u < J->hi; ++u) J->nbad += bad(f_of((uint32_t)u), J->e, y);
    return NULL;
}

static uint64_t xs(uint64_t *s) { *s ^= *s << 13; *s ^= *s >> 7; *s ^= *s << 17; return *s; }

int main(int argc, char **argv) {
    uint64_t nbad = 0;
    if (argc == 3 && !strcmp(argv[1], "all")) {
        const float e = strtof(argv[2], NULL);
        pthread_t th[8];
        job J[8];
        for (int i = 0; i < 8; ++i) {
            J[i].e = e;
            J[i].lo = (uint64_t)i << 29;
            J[i].hi = (uint64_t)(i + 1) << 29;
            J[i].nbad = 0;
            pthread_create(&th[i], NULL, run_all, &J[i]);
        }
        for (int i = 0; i < 8; ++i) {
            pthread_join(th[i], NULL);
            nbad += J[i].nbad;
        }
        printf("extent %.9g: %llu mismatches over all numerators\n", e, (unsigned long long)nbad);
    } else if (argc == 4 && !strcmp(argv[1], "random")) {
        const uint64_t n = strtoull(argv[2], NULL, 10);
        uint64_t s = strtoull(argv[3], NULL, 10) * 2654435761ull + 1;
        for (int k = 0; k < 64; ++k) {
            const float e = ldexpf(1.0f + (float)(xs(&s) >> 40) * 0x1p-24f, (int)(xs(&s) % 31) - 20);
            const float y = 1.0f / e;
            for (uint64_t i = 0; i < n; ++i) {
                /* numerators of the size a lookup sees: up to a few extents,
                   and down to 2^-100 with a log-uniform exponent */
                const uint32_t m = (uint32_t)(xs(&s) & 0x807fffffu);
                const int ex = (int)(xs(&s) % 140) - 100;
                const float a = ldexpf(f_of(m | 0x3f800000u), ex) * (ex > 20 ? 0.0f : 1.0f) +
                                (ex > 20 ? f_of(m | 0x3f800000u) * e * 8.0f : 0.0f);
                nbad += bad(a, e, y);
            }
        }
        printf("64 random extents x %llu numerators: %llu mismatches\n", (unsigned long long)n,
               (unsigned long long)nbad);
    } else if (argc == 3 && !strcmp(argv[1], "tiny")) {
        const float e = strtof(argv[2], NULL), y = 1.0f / e;
        for (uint32_t u = 1; u < u_of(0x1p-100f); u += 97) {
            nbad += differs(f_of(u), e, y);
            nbad += differs(f_of(u | 0x80000000u), e, y);
        }
        printf("extent %.9g: the correction differs from the division for %llu sampled numerators below 2^-100\n", e,
               (unsigned long long)nbad);
        return nbad ? 0 : 1;
    } else {
        fprintf(stderr, "usage: markstein_check all <e> | random <n> <seed> | tiny <e>\n");
        return 2;
    }
    return nbad ? 1 : 0;
}
