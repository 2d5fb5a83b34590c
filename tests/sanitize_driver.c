/* Exercises the CPU oracle (oracle/mpf_oracle.c, test infrastructure) on small cases for a build
 * with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py): every code path
 * the tests use - isotropic 2D and 3D grain growth, two phases with cubic anisotropy and a constant
 * driving force averaged over the interface (with the per-step switches), the triple-junction
 * VALUE model, energies, fractions, grain records, the time-step report and the standalone
 * limiter. Exit status 0 when every call returns ORC_OK. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpf_oracle.h"

static unsigned long long rng_state = 0x9E3779B97F4A7C15ull;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

#define CHECK(x)                                                        \
    do {                                                                \
        int rc_ = (x);                                                  \
        if (rc_ != ORC_OK) {                                            \
            fprintf(stderr, "%s:%d: %s -> %d\n", __FILE__, __LINE__, #x, rc_); \
            return 1;                                                   \
        }                                                               \
    } while (0)

/* Periodic Voronoi start: a cell holds its nearest seed's grain at value 1, or, within one cell
 * of the bisector of its two nearest seeds, both grains at 1/2 (the overlap band the step grows
 * interfaces from). count[ncells]; index / value concatenated in cell order. Returns the entries. */
static long voronoi(int nx, int ny, int nz, int ng, int32_t* count, int32_t* index, double* value)
{
    double sx[64], sy[64], sz[64];
    for (int g = 0; g < ng; ++g) {
        sx[g] = urand() * nx;
        sy[g] = urand() * ny;
        sz[g] = urand() * nz;
    }
    long c = 0, e = 0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z, ++c) {
                int b1 = 0, b2 = 0;
                double d1 = 1e300, d2 = 1e300;
                for (int g = 0; g < ng; ++g) {
                    double dx = fabs(x - sx[g]), dy = fabs(y - sy[g]), dz = fabs(z - sz[g]);
                    dx = fmin(dx, nx - dx);
                    dy = fmin(dy, ny - dy);
                    dz = fmin(dz, nz - dz);
                    const double d = sqrt(dx * dx + dy * dy + dz * dz);
                    if (d < d1) {
                        d2 = d1; b2 = b1;
                        d1 = d; b1 = g;
                    } else if (d < d2) {
                        d2 = d; b2 = g;
                    }
                }
                if (ng > 1 && d2 - d1 < 1.0) {
                    count[c] = 2;
                    index[e] = b1; value[e++] = 0.5;
                    index[e] = b2; value[e++] = 0.5;
                } else {
                    count[c] = 1;
                    index[e] = b1; value[e++] = 1.0;
                }
            }
    return e;
}

/* Runs one case: nsteps steps, then every query. */
static int run_case(const orc_params* p, int ng, const int32_t* phase, const double* rot, int nsteps, int toggles)
{
    const int nx = p->nx ? p->nx : 1, ny = p->ny ? p->ny : 1, nz = p->nz ? p->nz : 1;
    const long n = (long)nx * ny * nz;
    int32_t* count = malloc(sizeof(int32_t) * n);
    int32_t* index = malloc(sizeof(int32_t) * 2 * n);
    double* value = malloc(sizeof(double) * 2 * n);
    voronoi(nx, ny, nz, ng, count, index, value);
    orc_ctx* c = NULL;
    CHECK(orc_create(p, ng, phase, rot, NULL, &c));
    CHECK(orc_load(c, count, index, value));
    CHECK(orc_finalize_initialization(c));
    CHECK(orc_keep_step_start(c, 1));
    const double dt = 1e-4;  /* below every case's stability limit (1.67e-4 in 3D at these tables) */
    for (int s = 0; s < nsteps; ++s) {
        if (toggles) CHECK(orc_set_step_terms(c, s % 3 != 1, s % 2 == 0));
        CHECK(orc_step(c, dt, 1));
    }
    (void)orc_max_time_step(c);
    double e_prev = 0.0, e_cur = 0.0;
    CHECK(orc_interface_energy(c, ORC_IP_PREVIOUS, &e_prev));
    CHECK(orc_interface_energy(c, ORC_IP_CURRENT, &e_cur));
    double* frac = malloc(sizeof(double) * n * p->nphases);
    CHECK(orc_phase_fractions(c, frac));
    const long tot = (long)orc_total_entries(c);
    int32_t* dc = malloc(sizeof(int32_t) * n);
    int32_t* di = malloc(sizeof(int32_t) * (tot + 1));
    double* dv = malloc(sizeof(double) * (tot + 1));
    int8_t* fl = malloc(n);
    CHECK(orc_download(c, dc, di, dv, fl));
    double* vol = malloc(sizeof(double) * ng);
    double* vr = malloc(sizeof(double) * ng);
    double* mv = malloc(sizeof(double) * ng);
    double* rv = malloc(sizeof(double) * ng);
    uint8_t* ex = malloc(ng);
    int32_t* st = malloc(sizeof(int32_t) * ng);
    CHECK(orc_grains(c, vol, vr));
    CHECK(orc_grain_records(c, mv, rv, ex, st));
    int32_t mm = 0, mf = 0;
    orc_list_stats(c, &mm, &mf);
    printf("case %dx%dx%d: %ld entries, energy %.6e / %.6e, warnings %lld, lists %d / %d\n", p->nx, p->ny, p->nz,
           tot, e_prev, e_cur, (long long)orc_limiter_warnings(c), mm, mf);
    orc_destroy(c);
    free(count); free(index); free(value); free(frac); free(dc); free(di); free(dv); free(fl);
    free(vol); free(vr); free(mv); free(rv); free(ex); free(st);
    return 0;
}

int main(void)
{
    orc_set_threads(2);
    /* 1. isotropic 2D grain growth */
    {
        const double sigma[1] = {0.24}, mu[1] = {4e-9};
        orc_params p;
        memset(&p, 0, sizeof p);
        p.nx = 32; p.ny = 32; p.nz = 0; p.dx = 1e-6; p.iwidth = 5.0; p.nphases = 1;
        p.sigma = sigma; p.mu = mu; p.tj_model = ORC_TJ_MODEL; p.consider_nucleus_volume = 1;
        p.nucleus_volume_factor = 1.0; p.dg_limit = 0.95; p.regularization_factor = 1.0;
        const int32_t phase[8] = {0};
        if (run_case(&p, 8, phase, NULL, 30, 0)) return 1;
    }
    /* 2. two phases, cubic anisotropy, driving force averaged over the interface */
    {
        const double sigma[4] = {0.24, 0.3, 0.3, 0.2}, mu[4] = {4e-9, 3e-9, 3e-9, 5e-9};
        const int32_t em[4] = {ORC_EN_CUBIC, ORC_EN_CUBIC, ORC_EN_CUBIC, ORC_EN_CUBIC};
        const double eps[4] = {0.05, 0.05, 0.05, 0.05};
        const int32_t dgm[4] = {0, 1, 1, 0};
        const double dgv[4] = {0.0, 1e5, -1e5, 0.0};
        orc_params p;
        memset(&p, 0, sizeof p);
        p.nx = 24; p.ny = 24; p.nz = 0; p.dx = 1e-6; p.iwidth = 5.0; p.nphases = 2;
        p.sigma = sigma; p.mu = mu; p.energy_model = em; p.epsilon = eps; p.tj_model = ORC_TJ_MODEL;
        p.consider_nucleus_volume = 1; p.nucleus_volume_factor = 1.0; p.dg_mode = dgm; p.dg_value = dgv;
        p.dg_limit = 0.95; p.regularization_factor = 1.0; p.dg_average = 1; p.dg_weights = 0;
        int32_t phase[6];
        double rot[6 * 9];
        for (int g = 0; g < 6; ++g) {
            phase[g] = g % 2;
            /* rotation about z by a random angle */
            const double a = 6.283185307179586 * urand(), ca = cos(a), sa = sin(a);
            const double r[9] = {ca, -sa, 0.0, sa, ca, 0.0, 0.0, 0.0, 1.0};
            memcpy(rot + 9 * g, r, sizeof r);
        }
        if (run_case(&p, 6, phase, rot, 12, 1)) return 1;
    }
    /* 3. 3D, triple-junction VALUE model, RANGE weights */
    {
        const double sigma[1] = {0.24}, mu[1] = {4e-9};
        orc_params p;
        memset(&p, 0, sizeof p);
        p.nx = 12; p.ny = 12; p.nz = 12; p.dx = 1e-6; p.iwidth = 4.0; p.nphases = 1;
        p.sigma = sigma; p.mu = mu; p.tj_model = ORC_TJ_VALUE; p.tj_energy = 0.1;
        p.consider_nucleus_volume = 1; p.nucleus_volume_factor = 1.0; p.dg_limit = 0.95;
        p.regularization_factor = 1.0;
        const int32_t phase[5] = {0};
        if (run_case(&p, 5, phase, NULL, 6, 0)) return 1;
    }
    /* 4. the standalone limiter on random lists and increments */
    for (int t = 0; t < 2000; ++t) {
        const int n = 2 + (int)(urand() * 6);
        int32_t idx[8];
        double val[8], sum = 0.0;
        for (int i = 0; i < n; ++i) {
            idx[i] = i * 3 + 1;
            val[i] = urand() < 0.2 ? 0.0 : urand();
            sum += val[i];
        }
        for (int i = 0; i < n; ++i) val[i] = sum > 0.0 ? val[i] / sum : 1.0 / n;
        int32_t pa[32], pb[32];
        double v1[32];
        int32_t np = 0;
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n && np < 28; ++b) {
                pa[np] = idx[a];
                pb[np] = idx[b];
                v1[np] = (urand() - 0.5) * 2e4;
                ++np;
            }
        int64_t warnings = 0;
        CHECK(orc_normalize_pairs(n, idx, val, &np, pa, pb, v1, 1e-4, &warnings));
    }
    printf("sanitize driver ok\n");
    return 0;
}
