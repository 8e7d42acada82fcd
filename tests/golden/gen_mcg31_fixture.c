/* Writes golden vectors for the two MCG31 streams of the LPN task, produced by
 * Intel MKL's VSL (libmkl_rt.so, MKL 2021.4.0, the conda runtime of this image):
 *   - perturbation streams (init_streams_var, newdel_exp.cpp:313-330):
 *       vslNewStream(&s, VSL_BRNG_MCG31, 777); vslCopyStream; vslLeapfrogStream(s, tr, M);
 *       vdRngUniform(VSL_RNG_METHOD_UNIFORM_STD, s, 2 d, v, -1.0, 1.0)   (experiment.cpp:2311)
 *   - random-observable disorder (init_random_obs, newdel.cpp:3652-3690):
 *       vslNewStream(&s, VSL_BRNG_MCG31, 3873247); vslLeapfrogStream(s, seed + obs, mns);
 *       vdRngGaussian(VSL_RNG_METHOD_GAUSSIAN_BOXMULLER, s, n, v, 0.0, 1.0)
 *   - a plain [0, 1) stream for the integer-to-double conversion.
 * Build+run (this container only):  gcc -O1 gen_mcg31_fixture.c -ldl -o /tmp/gen31 && /tmp/gen31 > mcg31_mkl.json
 * The prototypes below are the documented VSL C interface; the numeric ids are
 * VSL_BRNG_MCG31 = 1<<20, VSL_RNG_METHOD_UNIFORM_STD = 0 and
 * VSL_RNG_METHOD_GAUSSIAN_BOXMULLER = 0. */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

typedef void *vsl_stream;

static void print_bits(const double *v, int n)
{
    for (int j = 0; j < n; j++) {
        uint64_t b;
        memcpy(&b, &v[j], 8);
        printf("\"%016llx\"%s", (unsigned long long)b, j + 1 < n ? ", " : "");
    }
}

int main(void)
{
    void *h = dlopen("/opt/conda/lib/libmkl_rt.so", RTLD_NOW);
    if (!h) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 1; }
    int (*new_stream)(vsl_stream *, int, unsigned) = (int (*)(vsl_stream *, int, unsigned))dlsym(h, "vslNewStream");
    int (*copy_stream)(vsl_stream *, vsl_stream) = (int (*)(vsl_stream *, vsl_stream))dlsym(h, "vslCopyStream");
    int (*leapfrog)(vsl_stream, int, int) = (int (*)(vsl_stream, int, int))dlsym(h, "vslLeapfrogStream");
    int (*uniform)(int, vsl_stream, int, double *, double, double) =
        (int (*)(int, vsl_stream, int, double *, double, double))dlsym(h, "vdRngUniform");
    int (*gaussian)(int, vsl_stream, int, double *, double, double) =
        (int (*)(int, vsl_stream, int, double *, double, double))dlsym(h, "vdRngGaussian");
    if (!new_stream || !copy_stream || !leapfrog || !uniform || !gaussian) { fprintf(stderr, "dlsym\n"); return 1; }
    const int MCG31 = 1 << 20;

    /* perturbation streams: seed 777, leapfrog(tr, M), uniform [-1, 1) */
    const int M = 100, trs[] = {0, 1, 2, 49, 50, 51, 99};
    const int ntr = (int)(sizeof(trs) / sizeof(trs[0])), nvar = 42;   /* 2 d at d = 21 */
    vsl_stream base;
    if (new_stream(&base, MCG31, 777)) return 1;
    printf("{\n  \"generator\": \"MKL VSL MCG31\",\n  \"mkl_version\": \"2021.4.0\",\n");
    printf("  \"var\": {\"seed\": 777, \"nstreams\": %d, \"a\": -1.0, \"b\": 1.0, \"streams\": [\n", M);
    for (int i = 0; i < ntr; i++) {
        vsl_stream s;
        double v[64];
        if (copy_stream(&s, base) || leapfrog(s, trs[i], M) || uniform(0, s, nvar, v, -1.0, 1.0)) return 1;
        printf("    {\"k\": %d, \"bits\": [", trs[i]);
        print_bits(v, nvar);
        printf("]}%s\n", i + 1 < ntr ? "," : "");
    }
    printf("  ]},\n");

    /* random-observable disorder: seed 3873247, leapfrog(seed + obs, mns), BoxMuller */
    const int mns = 1000000, ks[] = {1, 2, 5}, nk = 3, ng = 242;   /* 2 d^2 at d = 11 */
    vsl_stream gbase;
    if (new_stream(&gbase, MCG31, 3873247)) return 1;
    printf("  \"gauss\": {\"seed\": 3873247, \"mns\": %d, \"streams\": [\n", mns);
    for (int i = 0; i < nk; i++) {
        vsl_stream s;
        static double g[512];
        if (copy_stream(&s, gbase) || leapfrog(s, ks[i], mns) || gaussian(0, s, ng, g, 0.0, 1.0)) return 1;
        printf("    {\"k\": %d, \"bits\": [", ks[i]);
        print_bits(g, ng);
        printf("]}%s\n", i + 1 < nk ? "," : "");
    }
    printf("  ]},\n");

    /* plain stream, seed 777: [0, 1) */
    vsl_stream s;
    static double u[1024];
    if (copy_stream(&s, base) || uniform(0, s, 1024, u, 0.0, 1.0)) return 1;
    printf("  \"plain_1024\": [");
    print_bits(u, 1024);
    printf("]\n}\n");
    return 0;
}
