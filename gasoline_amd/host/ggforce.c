/*
 * ggforce -- one gravity step and one SPH step through the C ABI, driven
 * from C the way gasoline's master drives its per-process code.
 *
 *   ggforce [Ngrav] [Nsph] [fp64]
 *
 * Gravity: a Plummer sphere of Ngrav particles (G = M = a = 1, eps 0.01,
 * theta 0.7, quadrupole, bucket 16).  SPH: Nsph gas particles uniform in a
 * periodic unit box, u = 1, nSmooth 64.  Prints gasoline's log lines and
 * two conservation checks; exits non-zero on any failure.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gg_msr.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double urand(void)   /* xorshift64*, uniform in [0,1) */
{
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
}

static void iso(double r, double *x)
{
    double z = 2 * urand() - 1, ph = 2 * M_PI * urand(), s = sqrt(1 - z * z);
    x[0] = (float)(r * s * cos(ph));
    x[1] = (float)(r * s * sin(ph));
    x[2] = (float)(r * z);
}

int main(int argc, char **argv)
{
    int ng = argc > 1 ? atoi(argv[1]) : 65536;
    int ns = argc > 2 ? atoi(argv[2]) : 65536;
    int prec = (argc > 3 && strcmp(argv[3], "fp64") == 0) ? GG_PREC_FP64 : GG_PREC_FP32;
    GGPKD pkd;
    int i, rc = 0;

    /* ---- gravity */
    {
        double *r = malloc(sizeof(double) * 3 * ng), *m = malloc(sizeof(double) * ng);
        double *h = malloc(sizeof(double) * ng), *a = malloc(sizeof(double) * 3 * ng);
        double p[3] = {0, 0, 0}, amax = 0;
        ggGravityOut out;
        for (i = 0; i < ng; ++i) {
            double x1 = 0.999 * urand() + 1e-10;
            iso(1.0 / sqrt(pow(x1, -2.0 / 3.0) - 1.0), r + 3 * i);
            m[i] = 1.0 / ng;
            h[i] = 0.01;
        }
        if (gg_initialize(&pkd, 0, ng) || gg_set_precision(pkd, prec) ||
            gg_put(pkd, GG_R, r, 0) || gg_put(pkd, GG_MASS, m, 0) ||
            gg_put(pkd, GG_SOFT, h, 0)) {
            fprintf(stderr, "setup failed: %s\n", gg_last_error());
            return 1;
        }
        if (ggMsrGravity(pkd, 16, 0.7, 2, 1, &out)) return 1;
        gg_get(pkd, GG_A, a, 0);
        for (i = 0; i < ng; ++i) {
            int k;
            for (k = 0; k < 3; ++k) {
                p[k] += m[i] * a[3 * i + k];
                if (fabs(m[i] * a[3 * i + k]) > amax) amax = fabs(m[i] * a[3 * i + k]);
            }
        }
        printf("gravity: nActive %d, net force %.3e of max %.3e\n", out.nActive,
               sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]), amax);
        if (out.nActive != ng) rc = 1;
        gg_finish(pkd);
        free(r); free(m); free(h); free(a);
    }
    /* ---- SPH */
    {
        double *r = malloc(sizeof(double) * 3 * ns), *v = malloc(sizeof(double) * 3 * ns);
        double *m = malloc(sizeof(double) * ns), *u = malloc(sizeof(double) * ns);
        double *a = malloc(sizeof(double) * 3 * ns), *rho = malloc(sizeof(double) * ns);
        double L[3] = {1, 1, 1}, p[3] = {0, 0, 0}, pabs = 0, rmean = 0;
        gg_smf smf = {0.0, 1.0, 1.0, 2.0, 5.0 / 3.0, 0.4 * 2 / 1.6, 0.2, 0, 0, 0, 1, 0, 0, 0};
        for (i = 0; i < ns; ++i) {
            int k;
            for (k = 0; k < 3; ++k) {
                r[3 * i + k] = floor(urand() * 16777216.0) / 16777216.0;
                v[3 * i + k] = (float)(0.1 * sqrt(-2 * log(1 - urand())) * cos(2 * M_PI * urand()));
            }
            m[i] = 1.0 / ns;
            u[i] = 1.0;
        }
        if (gg_initialize(&pkd, 0, ns) || gg_set_precision(pkd, prec) ||
            gg_put(pkd, GG_R, r, 0) || gg_put(pkd, GG_VPRED, v, 0) ||
            gg_put(pkd, GG_MASS, m, 0) || gg_put(pkd, GG_UPRED, u, 0)) {
            fprintf(stderr, "setup failed: %s\n", gg_last_error());
            return 1;
        }
        if (ggMsrSph(pkd, &smf, 64, 1, L, 1)) return 1;
        gg_get(pkd, GG_A, a, 0);
        gg_get(pkd, GG_DENSITY, rho, 0);
        for (i = 0; i < ns; ++i) {
            int k;
            for (k = 0; k < 3; ++k) {
                p[k] += m[i] * a[3 * i + k];
                pabs += fabs(m[i] * a[3 * i + k]);
            }
            rmean += rho[i] / ns;
        }
        printf("sph: mean density %.4f, net momentum change %.3e of %.3e\n", rmean,
               sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]), pabs);
        if (!(fabs(p[0]) + fabs(p[1]) + fabs(p[2]) < 1e-6 * pabs)) rc = 1;
        if (!(rmean > 0.9 && rmean < 1.3)) rc = 1;
        gg_finish(pkd);
        free(r); free(v); free(m); free(u); free(a); free(rho);
    }
    printf(rc ? "ggforce FAILED\n" : "ggforce ok\n");
    return rc;
}
