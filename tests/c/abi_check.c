/*
 * abi_check.c -- the C ABI of libsmilei_amd.so exercised from C99.
 *
 * Built with gcc -std=c99 -pedantic -Werror against include/smilei_amd.h
 * and linked to the library (tests/test_c_abi.py).  Modes:
 *   symbols           takes the address of every declared entry point (the
 *                     link fails on a missing one); no device work
 *   context  DIR      whole-step API: create, species, particles, steps,
 *                     fields / particles / energies back
 *   ops3d    DIR      per-operator 3D entry points on one patch
 *   ops2d    DIR      per-operator 2D entry points on one patch
 *   seams    DIR      the two reference-seam deposits on a binned patch
 *   sync     DIR      J ghost sum, B exchange, sort/exchange over patches
 * Inputs are raw little-endian arrays DIR/<name>, written by the test;
 * outputs go to DIR/out_<name>; the test compares them with the oracle.
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "smilei_amd.h"

static const char *g_dir = ".";

static void die(const char *what, const char *detail)
{
    fprintf(stderr, "abi_check: %s: %s\n", what, detail ? detail : "");
    exit(1);
}

#define SMI(call)                                                  \
    do {                                                           \
        if ((call) != SMI_OK) die(#call, smi_last_error());        \
    } while (0)
#define HIPC(call)                                                 \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) die(#call, hipGetErrorString(e_));   \
    } while (0)

static void *load(const char *name, size_t *bytes)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open", path);
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    void *p = malloc(sz > 0 ? (size_t)sz : 1);
    if (!p || (sz > 0 && fread(p, 1, (size_t)sz, f) != (size_t)sz)) die("cannot read", path);
    fclose(f);
    if (bytes) *bytes = (size_t)sz;
    return p;
}

static void save(const char *name, const void *p, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/out_%s", g_dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || (bytes && fwrite(p, 1, bytes, f) != bytes)) die("cannot write", path);
    fclose(f);
}

/* host file -> new device buffer; *bytes receives its size */
static void *dload(const char *name, size_t *bytes)
{
    size_t b = 0;
    void *h = load(name, &b);
    void *d = NULL;
    HIPC(hipMalloc(&d, b > 0 ? b : 8));
    if (b) HIPC(hipMemcpy(d, h, b, hipMemcpyHostToDevice));
    free(h);
    if (bytes) *bytes = b;
    return d;
}

static void *dzeros(size_t bytes)
{
    void *d = NULL;
    void *h = calloc(bytes > 0 ? bytes : 8, 1);
    HIPC(hipMalloc(&d, bytes > 0 ? bytes : 8));
    HIPC(hipMemcpy(d, h, bytes > 0 ? bytes : 8, hipMemcpyHostToDevice));
    free(h);
    return d;
}

static void dsave(const char *name, const void *d, size_t bytes)
{
    void *h = malloc(bytes > 0 ? bytes : 8);
    HIPC(hipDeviceSynchronize());
    if (bytes) HIPC(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    save(name, h, bytes);
    free(h);
}

typedef void (*fn_t)(void);

static int mode_symbols(void)
{
    /* one entry per function the header declares (ISO C converts between
     * function pointer types, not to object pointers) */
    const fn_t fns[] = {
        (fn_t)smi_last_error, (fn_t)smi_build_hash, (fn_t)smi_create,
        (fn_t)smi_destroy, (fn_t)smi_add_species, (fn_t)smi_set_particles,
        (fn_t)smi_init_species, (fn_t)smi_init_species_mj, (fn_t)smi_init_species_like,
        (fn_t)smi_init_species_profile, (fn_t)smi_init_species_profile_mj,
        (fn_t)smi_num_particles, (fn_t)smi_get_particles, (fn_t)smi_patch_counts,
        (fn_t)smi_set_field, (fn_t)smi_get_field, (fn_t)smi_step,
        (fn_t)smi_synchronize, (fn_t)smi_set_time_frozen, (fn_t)smi_set_step_count,
        (fn_t)smi_step_count, (fn_t)smi_energies, (fn_t)smi_scalars, (fn_t)smi_rho,
        (fn_t)smi_gauss, (fn_t)smi_profile_step, (fn_t)smi_stream,
        (fn_t)smi_set_tuning, (fn_t)smi_get_tuning, (fn_t)smi_autotune,
        (fn_t)smi_interpolate_3d, (fn_t)smi_interpolate_2d, (fn_t)smi_push_boris,
        (fn_t)smi_currents_3d, (fn_t)smi_currents_2d, (fn_t)smi_maxwell_ampere_3d,
        (fn_t)smi_maxwell_faraday_3d, (fn_t)smi_density_3d, (fn_t)smi_density_2d,
        (fn_t)smi_maxwell_ampere_2d, (fn_t)smi_maxwell_faraday_2d,
        (fn_t)smi_current_deposition_3d, (fn_t)smi_density_deposition_3d,
        (fn_t)smi_sum_j_ghosts, (fn_t)smi_exchange_b, (fn_t)smi_sort_exchange,
        (fn_t)smi_sort_exchange_noalloc, (fn_t)smi_sort_exchange_noalloc_max_words,
    };
    size_t k;
    for (k = 0; k < sizeof fns / sizeof fns[0]; ++k)
        if (!fns[k]) die("null entry point", NULL);
    /* argument checks run before any device work */
    {
        const int z[3] = {0, 0, 0};
        const double one[3] = {1., 1., 1.};
        if (smi_interpolate_3d(3, 0, 0, 0, 0, 0, 0, z, z, one, 0, 0, 0, 0, 0, 0, 0, 0, 0) != SMI_ERR_INVALID)
            die("order 3 accepted", NULL);
    }
    printf("%u entry points, build %s\n", (unsigned)(sizeof fns / sizeof fns[0]), smi_build_hash());
    return 0;
}

/* geom.i32: ndim order n_cells[3] patch_size[3] deposit nsteps precision
 * cell.f64: dx dy dz dt; species.f64: mass charge (one species);
 * x y z px py pz w .f64 */
static int mode_context(void)
{
    int *gi = (int *)load("geom.i32", NULL);
    double *cl = (double *)load("cell.f64", NULL);
    double *spd = (double *)load("species.f64", NULL);
    size_t b = 0;
    double *arr[7];
    const char *names[7] = {"x.f64", "y.f64", "z.f64", "px.f64", "py.f64", "pz.f64", "w.f64"};
    int q, id = -1, f;
    int64_t n, np_;
    smi_params p;
    smi_context *ctx = NULL;
    double *buf;
    int32_t *pk, *kk;
    double en[2];
    for (q = 0; q < 7; ++q) arr[q] = (double *)load(names[q], &b);
    n = (int64_t)(b / sizeof(double));
    memset(&p, 0, sizeof p);
    p.geometry = gi[0] == 3 ? "3Dcartesian" : "2Dcartesian";
    p.interpolation_order = gi[1];
    for (q = 0; q < 3; ++q) {
        p.n_cells[q] = gi[2 + q];
        p.patch_size[q] = gi[5 + q];
        p.cell_length[q] = cl[q];
    }
    p.timestep = cl[3];
    p.deposit = gi[8];
    p.precision = gi[10];
    SMI(smi_create(&p, &ctx));
    SMI(smi_add_species(ctx, "electron", spd[0], spd[1], n, &id));
    SMI(smi_set_particles(ctx, id, n, arr[0], arr[1], arr[2], arr[3], arr[4], arr[5], arr[6]));
    SMI(smi_step(ctx, gi[9]));
    SMI(smi_synchronize(ctx));
    {
        const int64_t nc = (int64_t)p.n_cells[0] * p.n_cells[1] * (gi[0] == 3 ? p.n_cells[2] : 1);
        buf = (double *)malloc((size_t)nc * SMI_NFIELDS * sizeof(double));
        for (f = 0; f < SMI_NFIELDS; ++f) SMI(smi_get_field(ctx, f, buf + (size_t)f * nc));
        save("fields.f64", buf, (size_t)nc * SMI_NFIELDS * sizeof(double));
        free(buf);
    }
    np_ = smi_num_particles(ctx, id);
    buf = (double *)malloc((size_t)(np_ > 0 ? np_ : 1) * 7 * sizeof(double));
    pk = (int32_t *)malloc((size_t)(np_ > 0 ? np_ : 1) * sizeof(int32_t));
    kk = (int32_t *)malloc((size_t)(np_ > 0 ? np_ : 1) * sizeof(int32_t));
    SMI(smi_get_particles(ctx, id, buf, buf + np_, buf + 2 * np_, buf + 3 * np_, buf + 4 * np_, buf + 5 * np_,
                          buf + 6 * np_, pk, kk));
    save("parts.f64", buf, (size_t)np_ * 7 * sizeof(double));
    save("patch.i32", pk, (size_t)np_ * sizeof(int32_t));
    save("key.i32", kk, (size_t)np_ * sizeof(int32_t));
    SMI(smi_energies(ctx, en));
    save("energies.f64", en, sizeof en);
    SMI(smi_destroy(ctx));
    free(buf);
    free(pk);
    free(kk);
    for (q = 0; q < 7; ++q) free(arr[q]);
    free(gi);
    free(cl);
    free(spd);
    return 0;
}

/* geom.i32: order np[3] dom[3]; cell.f64: dx dy dz dt charge mass;
 * fields Ex Ey Ez Bxm Bym Bzm (.f64, ghosted patch arrays), x y z px py pz w,
 * then for Maxwell: E?/B?/J? .f64 (E? and B? updated in place) */
static int mode_ops3d(void)
{
    int *gi = (int *)load("geom.i32", NULL);
    double *cl = (double *)load("cell.f64", NULL);
    const int order = gi[0], *np3 = gi + 1, *dom = gi + 4;
    const char *fn[6] = {"Ex.f64", "Ey.f64", "Ez.f64", "Bxm.f64", "Bym.f64", "Bzm.f64"};
    const char *pn[7] = {"x.f64", "y.f64", "z.f64", "px.f64", "py.f64", "pz.f64", "w.f64"};
    double *F[6], *P[7], *Ep, *Bp, *dold, *invgf, *J[3], *rho;
    int32_t *iold;
    size_t b = 0;
    int q;
    int64_t n;
    for (q = 0; q < 6; ++q) F[q] = (double *)dload(fn[q], NULL);
    for (q = 0; q < 7; ++q) P[q] = (double *)dload(pn[q], &b);
    n = (int64_t)(b / sizeof(double));
    Ep = (double *)dzeros((size_t)n * 3 * sizeof(double));
    Bp = (double *)dzeros((size_t)n * 3 * sizeof(double));
    dold = (double *)dzeros((size_t)n * 3 * sizeof(double));
    invgf = (double *)dzeros((size_t)n * sizeof(double));
    iold = (int32_t *)dzeros((size_t)n * 3 * sizeof(int32_t));
    SMI(smi_interpolate_3d(order, F[0], F[1], F[2], F[3], F[4], F[5], np3, dom, cl, n, P[0], P[1], P[2], Ep, Bp, iold,
                           dold, NULL));
    dsave("Epart.f64", Ep, (size_t)n * 3 * sizeof(double));
    dsave("Bpart.f64", Bp, (size_t)n * 3 * sizeof(double));
    dsave("iold.i32", iold, (size_t)n * 3 * sizeof(int32_t));
    dsave("deltaold.f64", dold, (size_t)n * 3 * sizeof(double));
    SMI(smi_push_boris(3, cl[4], cl[5], cl[3], n, P[0], P[1], P[2], P[3], P[4], P[5], Ep, Bp, invgf, NULL));
    for (q = 0; q < 6; ++q) dsave(pn[q], P[q], (size_t)n * sizeof(double));
    dsave("invgf.f64", invgf, (size_t)n * sizeof(double));
    {
        const size_t sz[3] = {(size_t)(np3[0] + 1) * np3[1] * np3[2], (size_t)np3[0] * (np3[1] + 1) * np3[2],
                              (size_t)np3[0] * np3[1] * (np3[2] + 1)};
        for (q = 0; q < 3; ++q) J[q] = (double *)dzeros(sz[q] * sizeof(double));
        SMI(smi_currents_3d(order, J[0], J[1], J[2], np3, dom, cl, cl[3], cl[4], n, P[0], P[1], P[2], P[6], iold,
                            dold, 1, NULL));
        dsave("Jx.f64", J[0], sz[0] * sizeof(double));
        dsave("Jy.f64", J[1], sz[1] * sizeof(double));
        dsave("Jz.f64", J[2], sz[2] * sizeof(double));
        rho = (double *)dzeros((size_t)np3[0] * np3[1] * np3[2] * sizeof(double));
        SMI(smi_density_3d(order, rho, np3, dom, cl, cl[4], n, P[0], P[1], P[2], P[6], NULL));
        dsave("rho.f64", rho, (size_t)np3[0] * np3[1] * np3[2] * sizeof(double));
    }
    {
        /* Maxwell on independent random arrays (E, B, J of the patch) */
        const char *en[3] = {"mE0.f64", "mE1.f64", "mE2.f64"}, *bn[3] = {"mB0.f64", "mB1.f64", "mB2.f64"};
        const char *jn[3] = {"mJ0.f64", "mJ1.f64", "mJ2.f64"};
        double *E[3], *B[3], *MJ[3];
        size_t eb[3], bb[3];
        for (q = 0; q < 3; ++q) {
            E[q] = (double *)dload(en[q], &eb[q]);
            B[q] = (double *)dload(bn[q], &bb[q]);
            MJ[q] = (double *)dload(jn[q], NULL);
        }
        SMI(smi_maxwell_ampere_3d(E[0], E[1], E[2], B[0], B[1], B[2], MJ[0], MJ[1], MJ[2], np3, cl[3], cl, NULL));
        SMI(smi_maxwell_faraday_3d(B[0], B[1], B[2], E[0], E[1], E[2], np3, cl[3], cl, NULL));
        for (q = 0; q < 3; ++q) {
            dsave(en[q], E[q], eb[q]);
            dsave(bn[q], B[q], bb[q]);
        }
    }
    free(gi);
    free(cl);
    return 0;
}

/* geom.i32: np[2] dom[2]; cell.f64: dx dy dt charge mass; fields as ops3d
 * (2D shapes), x y px py pz w; Maxwell arrays mE?/mB?/mJ? */
static int mode_ops2d(void)
{
    int *gi = (int *)load("geom.i32", NULL);
    double *cl = (double *)load("cell.f64", NULL);
    const int *np2 = gi, *dom = gi + 2;
    const char *fn[6] = {"Ex.f64", "Ey.f64", "Ez.f64", "Bxm.f64", "Bym.f64", "Bzm.f64"};
    const char *pn[6] = {"x.f64", "y.f64", "px.f64", "py.f64", "pz.f64", "w.f64"};
    double *F[6], *P[6], *Ep, *Bp, *dold, *invgf, *J[3], *rho, zero = 0.;
    int32_t *iold;
    size_t b = 0;
    int q;
    int64_t n;
    for (q = 0; q < 6; ++q) F[q] = (double *)dload(fn[q], NULL);
    for (q = 0; q < 6; ++q) P[q] = (double *)dload(pn[q], &b);
    n = (int64_t)(b / sizeof(double));
    Ep = (double *)dzeros((size_t)n * 3 * sizeof(double));
    Bp = (double *)dzeros((size_t)n * 3 * sizeof(double));
    dold = (double *)dzeros((size_t)n * 2 * sizeof(double));
    invgf = (double *)dzeros((size_t)n * sizeof(double));
    iold = (int32_t *)dzeros((size_t)n * 2 * sizeof(int32_t));
    (void)zero;
    SMI(smi_interpolate_2d(2, F[0], F[1], F[2], F[3], F[4], F[5], np2, dom, cl, n, P[0], P[1], Ep, Bp, iold, dold,
                           NULL));
    dsave("Epart.f64", Ep, (size_t)n * 3 * sizeof(double));
    dsave("Bpart.f64", Bp, (size_t)n * 3 * sizeof(double));
    dsave("iold.i32", iold, (size_t)n * 2 * sizeof(int32_t));
    dsave("deltaold.f64", dold, (size_t)n * 2 * sizeof(double));
    SMI(smi_push_boris(2, cl[3], cl[4], cl[2], n, P[0], P[1], NULL, P[2], P[3], P[4], Ep, Bp, invgf, NULL));
    for (q = 0; q < 5; ++q) dsave(pn[q], P[q], (size_t)n * sizeof(double));
    dsave("invgf.f64", invgf, (size_t)n * sizeof(double));
    {
        const size_t sz[3] = {(size_t)(np2[0] + 1) * np2[1], (size_t)np2[0] * (np2[1] + 1), (size_t)np2[0] * np2[1]};
        for (q = 0; q < 3; ++q) J[q] = (double *)dzeros(sz[q] * sizeof(double));
        SMI(smi_currents_2d(2, J[0], J[1], J[2], np2, dom, cl, cl[2], cl[3], n, P[0], P[1], P[4], P[5], invgf, iold,
                            dold, 1, NULL));
        dsave("Jx.f64", J[0], sz[0] * sizeof(double));
        dsave("Jy.f64", J[1], sz[1] * sizeof(double));
        dsave("Jz.f64", J[2], sz[2] * sizeof(double));
        rho = (double *)dzeros(sz[2] * sizeof(double));
        SMI(smi_density_2d(2, rho, np2, dom, cl, cl[3], n, P[0], P[1], P[5], NULL));
        dsave("rho.f64", rho, sz[2] * sizeof(double));
    }
    {
        const char *en[3] = {"mE0.f64", "mE1.f64", "mE2.f64"}, *bn[3] = {"mB0.f64", "mB1.f64", "mB2.f64"};
        const char *jn[3] = {"mJ0.f64", "mJ1.f64", "mJ2.f64"};
        double *E[3], *B[3], *MJ[3];
        size_t eb[3], bb[3];
        for (q = 0; q < 3; ++q) {
            E[q] = (double *)dload(en[q], &eb[q]);
            B[q] = (double *)dload(bn[q], &bb[q]);
            MJ[q] = (double *)dload(jn[q], NULL);
        }
        SMI(smi_maxwell_ampere_2d(E[0], E[1], E[2], B[0], B[1], B[2], MJ[0], MJ[1], MJ[2], np2, cl[2], cl, NULL));
        SMI(smi_maxwell_faraday_2d(B[0], B[1], B[2], E[0], E[1], E[2], np2, cl[2], cl, NULL));
        for (q = 0; q < 3; ++q) {
            dsave(en[q], E[q], eb[q]);
            dsave(bn[q], B[q], bb[q]);
        }
    }
    free(gi);
    free(cl);
    return 0;
}

/* geom.i32: nb np[3] dom[3] sizes Jx Jy Jz rho; cell.f64: icv dinv[3]
 * dxdt[3]; x y z w .f64, q.i16, bins.i32, iold.i32, deltaold.f64 */
static int mode_seams(void)
{
    int *gi = (int *)load("geom.i32", NULL);
    double *cl = (double *)load("cell.f64", NULL);
    const unsigned nb = (unsigned)gi[0];
    const int *np3 = gi + 1, *dom = gi + 4, *sz = gi + 7;
    size_t b = 0;
    double *x = (double *)dload("x.f64", &b), *y = (double *)dload("y.f64", NULL), *z = (double *)dload("z.f64", NULL);
    double *w = (double *)dload("w.f64", NULL), *dold = (double *)dload("deltaold.f64", NULL);
    short *qq = (short *)dload("q.i16", NULL);
    int *bins = (int *)dload("bins.i32", NULL), *iold = (int *)dload("iold.i32", NULL);
    const unsigned n = (unsigned)(b / sizeof(double));
    double *J[3], *rho;
    int c;
    for (c = 0; c < 3; ++c) J[c] = (double *)dzeros((size_t)sz[c] * sizeof(double));
    rho = (double *)dzeros((size_t)sz[3] * sizeof(double));
    SMI(smi_current_deposition_3d(J[0], J[1], J[2], sz[0], sz[1], sz[2], x, y, z, qq, w, bins, nb, nb, nb, NULL, iold,
                                  dold, n, cl[0], cl[1], cl[2], cl[3], cl[4], cl[5], cl[6], dom[0], dom[1], dom[2],
                                  np3[1], np3[2], 1, true));
    SMI(smi_density_deposition_3d(rho, sz[3], x, y, z, qq, w, bins, nb, nb, nb, NULL, iold, dold, n, cl[0], cl[1],
                                  cl[2], cl[3], cl[4], cl[5], cl[6], dom[0], dom[1], dom[2], np3[1], np3[2], 1, false));
    dsave("Jx.f64", J[0], (size_t)sz[0] * sizeof(double));
    dsave("Jy.f64", J[1], (size_t)sz[1] * sizeof(double));
    dsave("Jz.f64", J[2], (size_t)sz[2] * sizeof(double));
    dsave("rho.f64", rho, (size_t)sz[3] * sizeof(double));
    free(gi);
    free(cl);
    return 0;
}

/* geom.i32: ndim order n_cells[3] patch_size[3]; cell.f64: dx dy dz;
 * J.f64, B.f64 (patch-major blocks); x y z px py pz w .f64, q.i16 */
static int mode_sync(void)
{
    int *gi = (int *)load("geom.i32", NULL);
    double *cl = (double *)load("cell.f64", NULL);
    const int nd = gi[0], order = gi[1], *nc = gi + 2, *ps = gi + 5;
    size_t jb = 0, bb = 0, b = 0;
    double *J = (double *)dload("J.f64", &jb), *B = (double *)dload("B.f64", &bb);
    const char *pn[7] = {"x.f64", "y.f64", "z.f64", "px.f64", "py.f64", "pz.f64", "w.f64"};
    double *P[7];
    short *qq;
    int32_t *key, *cnt, *bins;
    int q, npatch = 1, clpp = 1;
    int64_t n;
    SMI(smi_sum_j_ghosts(nd, order, nc, ps, J, NULL));
    SMI(smi_exchange_b(nd, order, nc, ps, B, NULL));
    dsave("J.f64", J, jb);
    dsave("B.f64", B, bb);
    for (q = 0; q < 7; ++q) P[q] = (double *)dload(pn[q], q == 0 ? &b : NULL);
    qq = (short *)dload("q.i16", NULL);
    n = (int64_t)(b / sizeof(double));
    for (q = 0; q < nd; ++q) {
        npatch *= nc[q] / ps[q];
        clpp *= ps[q] / 4;
    }
    key = (int32_t *)dzeros((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    cnt = (int32_t *)dzeros((size_t)npatch * sizeof(int32_t));
    bins = (int32_t *)dzeros((size_t)npatch * clpp * sizeof(int32_t));
    SMI(smi_sort_exchange(nd, nc, ps, cl, n, P[0], P[1], nd == 3 ? P[2] : NULL, P[3], P[4], P[5], P[6], qq, key, cnt,
                          bins, NULL));
    for (q = 0; q < 7; ++q) dsave(pn[q], P[q], (size_t)n * sizeof(double));
    dsave("q.i16", qq, (size_t)n * sizeof(short));
    dsave("key.i32", key, (size_t)n * sizeof(int32_t));
    dsave("counts.i32", cnt, (size_t)npatch * sizeof(int32_t));
    dsave("bins.i32", bins, (size_t)npatch * clpp * sizeof(int32_t));
    free(gi);
    free(cl);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) die("usage", "abi_check symbols | context|ops3d|ops2d|seams|sync DIR");
    if (argc > 2) g_dir = argv[2];
    if (!strcmp(argv[1], "symbols")) return mode_symbols();
    if (!strcmp(argv[1], "context")) return mode_context();
    if (!strcmp(argv[1], "ops3d")) return mode_ops3d();
    if (!strcmp(argv[1], "ops2d")) return mode_ops2d();
    if (!strcmp(argv[1], "seams")) return mode_seams();
    if (!strcmp(argv[1], "sync")) return mode_sync();
    die("unknown mode", argv[1]);
    return 1;
}
