/* Writes golden vectors for the MCG59 leap-frog uniform stream used by the
 * trajectory engine, produced by Intel MKL's VSL (libmkl_rt.so, MKL 2021.4.0,
 * the conda runtime of this image).  The engine's per-trajectory stream is
 *   vslNewStream(&s, VSL_BRNG_MCG59, 777); vslCopyStream; vslLeapfrogStream(s, seed + tr, mns);
 *   vdRngUniform(VSL_RNG_METHOD_UNIFORM_STD, s, 1, &u, 0.0, 1.0)
 * (QJX newdel_exp.cpp:244-296, experiment.cpp:3261-3265, qj_proc.cpp:310-314).
 * Build+run (this container only):  gcc -O1 gen_mcg59_fixture.c -ldl -o /tmp/gen && /tmp/gen > mcg59_mkl.json
 * The prototypes below are the documented VSL C interface; the numeric ids are
 * VSL_BRNG_MCG59 = 4<<20 and VSL_RNG_METHOD_UNIFORM_STD = 0. */
#include <stdio.h>
#include <string.h>
#include <stdint.h>
#include <dlfcn.h>

typedef void *vsl_stream;

int main(void)
{
    void *h = dlopen("/opt/conda/lib/libmkl_rt.so", RTLD_NOW);
    if (!h) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 1; }
    int (*new_stream)(vsl_stream *, int, unsigned) = (int (*)(vsl_stream *, int, unsigned))dlsym(h, "vslNewStream");
    int (*copy_stream)(vsl_stream *, vsl_stream) = (int (*)(vsl_stream *, vsl_stream))dlsym(h, "vslCopyStream");
    int (*leapfrog)(vsl_stream, int, int) = (int (*)(vsl_stream, int, int))dlsym(h, "vslLeapfrogStream");
    int (*uniform)(int, vsl_stream, int, double *, double, double) =
        (int (*)(int, vsl_stream, int, double *, double, double))dlsym(h, "vdRngUniform");
    if (!new_stream || !copy_stream || !leapfrog || !uniform) { fprintf(stderr, "dlsym failed\n"); return 1; }

    const int ks[] = {0, 1, 2, 3, 17, 999, 123456, 999999};
    const int nk = (int)(sizeof(ks) / sizeof(ks[0]));
    const int mns = 1000000, ndraw = 64;
    vsl_stream base;
    if (new_stream(&base, 4 << 20, 777)) return 1;
    printf("{\n  \"generator\": \"MKL VSL MCG59 seed 777, leapfrog(k, 1000000), vdRngUniform STD [0,1)\",\n");
    printf("  \"mkl_version\": \"2021.4.0\",\n  \"mns\": %d,\n  \"streams\": [\n", mns);
    for (int i = 0; i < nk; i++) {
        vsl_stream s;
        if (copy_stream(&s, base)) return 1;
        if (leapfrog(s, ks[i], mns)) return 1;
        double u[64];
        if (uniform(0, s, ndraw, u, 0.0, 1.0)) return 1;
        printf("    {\"k\": %d, \"bits\": [", ks[i]);
        for (int j = 0; j < ndraw; j++) {
            uint64_t b; memcpy(&b, &u[j], 8);
            printf("\"%016llx\"%s", (unsigned long long)b, j + 1 < ndraw ? ", " : "");
        }
        printf("]}%s\n", i + 1 < nk ? "," : "");
    }
    /* plain (non leap-frog) stream: 4096 consecutive draws, for the conversion check */
    vsl_stream s;
    if (copy_stream(&s, base)) return 1;
    static double u[4096];
    if (uniform(0, s, 4096, u, 0.0, 1.0)) return 1;
    printf("  ],\n  \"plain_4096\": [");
    for (int j = 0; j < 4096; j++) {
        uint64_t b; memcpy(&b, &u[j], 8);
        printf("\"%016llx\"%s", (unsigned long long)b, j + 1 < 4096 ? ", " : "");
    }
    printf("]\n}\n");
    return 0;
}
