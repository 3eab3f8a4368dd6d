/*
 * TEST INFRASTRUCTURE ONLY — drives every entry point of the CPU oracle
 * (bfvm_oracle.c) on small, odd and degenerate shapes so that a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py,
 * `make -C oracle sanitize`) sees every loop bound and index: axes of 1 and 2
 * nodes, non-cubic grids, many substrates, Dirichlet nodes with and without
 * a substrate mask (repeated voxels included), bulk fields, agents stacked on
 * one voxel, an agent outside the domain (refused), the 2-D solver and the
 * gradient field.  Exits 0 when every call returns; a sanitizer report aborts
 * with a non-zero status.  The results are checked elsewhere (the oracle's
 * own tests); here only that they are finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define ORACLE_BULK 1
#define ORACLE_CELLS 2
#define ORACLE_DIFFUSION 4
#define ORACLE_LOD_2D 8

void oracle_thomas_coefficients(int n, int S, double dx, double dt, const double* D, const double* lambda,
                                double* denom, double* c);
void oracle_lod_step(int nx, int ny, int nz, int S, double dx, double dt, const double* D, const double* lambda,
                     double* rho, int n_dir, const int* dir_idx, const double* dir_val,
                     const unsigned char* dir_mask);
void oracle_bulk_step(size_t nvox, int S, double dt, double* rho, const double* supply, const double* target,
                      const double* uptake);
int oracle_nearest_voxel(int nx, int ny, int nz, double dx, double x0, double y0, double z0, double x, double y,
                         double z);
void oracle_agent_constants(int n_agents, int S, double dt, double voxel_volume, const double* volume,
                            const double* secretion, const double* saturation, const double* uptake,
                            double* temp1, double* temp2);
void oracle_agent_step(int S, double* rho, int n_agents, const int* voxel, const double* temp1,
                       const double* temp2);
int oracle_run(int nx, int ny, int nz, int S, double dx, const double* D, const double* lambda, double dt,
               int steps, int phases, double* rho, int n_dir, const int* dir_idx, const double* dir_val,
               const unsigned char* dir_mask, const double* bulk_supply, const double* bulk_target,
               const double* bulk_uptake, int n_agents, const double* agent_pos, const double* agent_volume,
               const double* agent_secretion, const double* agent_saturation, const double* agent_uptake,
               int* out_voxel);
void oracle_gradients(int nx, int ny, int nz, int S, double dx, double dy, double dz, const double* rho,
                      double* out);

static unsigned long long rng_state = 12345;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) / 9007199254740992.0;
}

static double* alloc_fill(size_t n, double lo, double hi)
{
    double* p = malloc(sizeof(double) * (n ? n : 1));
    if (!p) exit(3);
    for (size_t i = 0; i < n; ++i) p[i] = lo + (hi - lo) * urand();
    return p;
}

static int check_finite(const double* p, size_t n, const char* what)
{
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(p[i])) {
            fprintf(stderr, "%s: value %zu not finite\n", what, i);
            return 1;
        }
    return 0;
}

/* One grid through every phase and entry point. */
static int one_case(int nx, int ny, int nz, int S, int with_mask, int lod2d)
{
    const double dx = 10.0, dt = 0.01;
    const size_t nvox = (size_t)nx * ny * nz, nvs = nvox * (size_t)S;
    double* D = alloc_fill((size_t)S, 1e2, 1e5);
    double* lam = alloc_fill((size_t)S, 1e-3, 1.0);
    double* rho = alloc_fill(nvs, 0.0, 1.0);
    int bad = 0;

    /* Thomas coefficients of each axis length */
    const int ns[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        double* denom = malloc(sizeof(double) * (size_t)ns[a] * S);
        double* c = malloc(sizeof(double) * (size_t)ns[a] * S);
        oracle_thomas_coefficients(ns[a], S, dx, dt, D, lam, denom, c);
        bad |= check_finite(denom, (size_t)ns[a] * S, "denom");
        free(denom);
        free(c);
    }

    /* Dirichlet nodes: the first and last voxel, a repeated one, a middle one */
    const int n_dir = 4;
    int dir_idx[4] = {0, (int)(nvox - 1), 0, (int)(nvox / 2)};
    double* dir_val = alloc_fill((size_t)n_dir * S, 0.0, 1.0);
    unsigned char* mask = malloc((size_t)S);
    for (int s = 0; s < S; ++s) mask[s] = (unsigned char)(s % 2 == 0);
    oracle_lod_step(nx, ny, nz, S, dx, dt, D, lam, rho, n_dir, dir_idx, dir_val, with_mask ? mask : NULL);
    bad |= check_finite(rho, nvs, "lod step");

    /* bulk */
    double* sup = alloc_fill(nvs, 0.0, 0.1);
    double* tgt = alloc_fill(nvs, 0.5, 1.0);
    double* upt = alloc_fill(nvs, 0.0, 0.1);
    oracle_bulk_step(nvox, S, dt, rho, sup, tgt, upt);
    bad |= check_finite(rho, nvs, "bulk step");

    /* agents: three stacked on voxel 0, the rest anywhere in the box (the
       upper faces included), their constants and one step */
    const int na = 9;
    double* pos = malloc(sizeof(double) * 3 * na);
    for (int a = 0; a < na; ++a) {
        pos[3 * a + 0] = a < 3 ? 0.5 * dx : urand() * nx * dx;
        pos[3 * a + 1] = a < 3 ? 0.5 * dx : urand() * ny * dx;
        pos[3 * a + 2] = a < 3 ? 0.5 * dx : urand() * nz * dx;
    }
    pos[3 * (na - 1) + 0] = nx * dx;  /* exactly on the upper face: clamped to the last voxel */
    double* vol = alloc_fill((size_t)na, 1000.0, 3000.0);
    double* sec = alloc_fill((size_t)na * S, 0.0, 10.0);
    double* sat = alloc_fill((size_t)na * S, 0.0, 1.0);
    double* aup = alloc_fill((size_t)na * S, 0.0, 10.0);
    int* voxel = malloc(sizeof(int) * na);
    for (int a = 0; a < na; ++a) {
        voxel[a] = oracle_nearest_voxel(nx, ny, nz, dx, 0.0, 0.0, 0.0, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
        if (voxel[a] < 0 || (size_t)voxel[a] >= nvox) {
            fprintf(stderr, "agent %d in the box got voxel %d\n", a, voxel[a]);
            bad = 1;
            voxel[a] = 0;
        }
    }
    if (oracle_nearest_voxel(nx, ny, nz, dx, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0) >= 0) {
        fprintf(stderr, "a position outside the box got a voxel\n");
        bad = 1;
    }
    double* t1 = malloc(sizeof(double) * (size_t)na * S);
    double* t2 = malloc(sizeof(double) * (size_t)na * S);
    oracle_agent_constants(na, S, dt, dx * dx * dx, vol, sec, sat, aup, t1, t2);
    oracle_agent_step(S, rho, na, voxel, t1, t2);
    bad |= check_finite(rho, nvs, "agent step");

    /* whole runs: every phase, with and without the 2-D solver */
    int* out_voxel = malloc(sizeof(int) * na);
    const int phases = ORACLE_BULK | ORACLE_CELLS | ORACLE_DIFFUSION | (lod2d ? ORACLE_LOD_2D : 0);
    int rc = oracle_run(nx, ny, nz, S, dx, D, lam, dt, 3, phases, rho, n_dir, dir_idx, dir_val,
                        with_mask ? mask : NULL, sup, tgt, upt, na, pos, vol, sec, sat, aup, out_voxel);
    if (rc != 0) {
        fprintf(stderr, "oracle_run returned %d\n", rc);
        bad = 1;
    }
    bad |= check_finite(rho, nvs, "run");
    /* an agent outside the domain: refused (-2), nothing stepped */
    pos[0] = -5.0 * dx;
    rc = oracle_run(nx, ny, nz, S, dx, D, lam, dt, 1, phases, rho, 0, NULL, NULL, NULL, NULL, NULL, NULL, na, pos,
                    vol, sec, sat, aup, out_voxel);
    if (rc != -2) {
        fprintf(stderr, "an agent outside the domain gave %d, not -2\n", rc);
        bad = 1;
    }

    /* gradients */
    double* g = malloc(sizeof(double) * nvs * 3);
    oracle_gradients(nx, ny, nz, S, dx, dx, dx, rho, g);
    bad |= check_finite(g, nvs * 3, "gradients");

    free(D); free(lam); free(rho); free(dir_val); free(mask); free(sup); free(tgt); free(upt);
    free(pos); free(vol); free(sec); free(sat); free(aup); free(voxel); free(t1); free(t2);
    free(out_voxel); free(g);
    return bad;
}

int main(void)
{
    static const int shapes[][4] = {
        {9, 7, 5, 3}, {1, 1, 6, 2}, {6, 1, 1, 1}, {1, 5, 1, 4}, {2, 2, 2, 1},
        {13, 11, 1, 5}, {4, 3, 9, 16}, {1, 1, 1, 3}, {17, 2, 3, 7},
    };
    int bad = 0, cases = 0;
    for (size_t k = 0; k < sizeof(shapes) / sizeof(shapes[0]); ++k)
        for (int with_mask = 0; with_mask < 2; ++with_mask) {
            const int lod2d = shapes[k][2] == 1;
            bad |= one_case(shapes[k][0], shapes[k][1], shapes[k][2], shapes[k][3], with_mask, lod2d);
            ++cases;
        }
    printf("%d cases, %s\n", cases, bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
