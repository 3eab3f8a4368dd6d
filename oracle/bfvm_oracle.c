/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle for the diffusion-decay path.
 *
 * A plain-C restatement of BioFVM 1.0.4's hot path, written to follow the
 * reference's arithmetic order operation by operation so that, compiled with
 * -ffp-contract=off (as the reference's ISO C++11 build is), it reproduces the
 * reference bit for bit.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker; the product path
 * never calls it.
 *
 * Layout: one flat array rho[((k*ny + j)*nx + i)*S + s] (voxel-major,
 * substrate innermost), which is the reference's vector<vector<double>>
 * flattened (BioFVM_microenvironment.h:51,63; voxel index BioFVM_mesh.cpp:641-644).
 *
 * Pinning (see tests/test_oracle_*.py): every function here is checked
 *   - bit for bit against the compiled reference (oracle/_ref) and against the
 *     committed golden fixtures generated from it (tests/golden/),
 *   - against closed forms: discrete cosine modes of the LOD step, exact mass
 *     decay (1+dt*lambda/3)^-3 per step, pure scaling on n==1 axes, dense
 *     solves of each tridiagonal line, and hand-composed bulk/agent updates.
 * The per-substrate Dirichlet mask is an extension the reference lacks; with
 * a NULL mask the behaviour is the reference's (whole vector overwritten).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_BULK 1
#define ORACLE_CELLS 2
#define ORACLE_DIFFUSION 4
#define ORACLE_LOD_2D 8 /* with ORACLE_DIFFUSION: the 2-D solver instead of the 3-D one */

/*
 * Thomas coefficients for one axis of n nodes, all S substrates
 * (BioFVM_solvers.cpp:104-142; the y and z axes repeat it at :144-170).
 * denom[i*S+s], c[i*S+s]; i in [0,n).
 */
static void thomas_coefficients_split(int n, int S, double dx, double dt,
                                      const double* D, const double* lambda, int dims,
                                      double* denom, double* c)
{
    for (int s = 0; s < S; ++s) {
        double c1 = D[s];
        c1 *= dt;
        c1 /= dx;
        c1 /= dx;                       /* dt*D/dx^2          :111-113 */
        double c1a = c1 * -1.0;         /* -dt*D/dx^2         :115-116 */
        double c2 = lambda[s];
        c2 *= dt;
        if (dims == 2) c2 *= 0.5;       /* dt*lambda/2, LOD_2D :313-314 */
        else c2 /= 3.0;                 /* dt*lambda/3        :118-119 */
        double c3 = 1.0;
        c3 += c1;
        c3 += c1;
        c3 += c2;                       /* 1 + 2c1 + c2       :121-123 */
        double c3a = 1.0;
        c3a += c1;
        c3a += c2;                      /* 1 + c1 + c2        :125-126 */

        for (int i = 0; i < n; ++i) {
            c[i * S + s] = c1a;
            denom[i * S + s] = c3;
        }
        denom[0 * S + s] = c3a;
        denom[(n - 1) * S + s] = c3a;
        if (n == 1) {
            denom[s] = 1.0;
            denom[s] += c2;             /* :134-135 */
        }
        c[s] /= denom[s];
        for (int i = 1; i <= n - 1; ++i) {
            denom[i * S + s] += c1 * c[(i - 1) * S + s];   /* axpy :139 */
            c[i * S + s] /= denom[i * S + s];
        }
    }
}

void oracle_thomas_coefficients(int n, int S, double dx, double dt,
                                const double* D, const double* lambda,
                                double* denom, double* c)
{
    thomas_coefficients_split(n, S, dx, dt, D, lambda, 3, denom, c);
}

/* c1[s] = dt*D/dx^2 in the reference's order (BioFVM_solvers.cpp:111-113). */
static void thomas_constant1(int S, double dx, double dt, const double* D, double* c1)
{
    for (int s = 0; s < S; ++s) {
        double v = D[s];
        v *= dt;
        v /= dx;
        v /= dx;
        c1[s] = v;
    }
}

/*
 * Dirichlet overwrite (BioFVM_microenvironment.cpp:192-197): every node's
 * whole substrate vector, or only the substrates with mask[s] != 0.
 */
void oracle_apply_dirichlet(int S, double* rho, int n_dir, const int* idx,
                            const double* val, const unsigned char* mask)
{
    for (int d = 0; d < n_dir; ++d) {
        double* r = rho + (size_t)idx[d] * S;
        for (int s = 0; s < S; ++s)
            if (!mask || mask[s]) r[s] = val[(size_t)d * S + s];
    }
}

/*
 * One Thomas line: n nodes, stride `jump` voxels, all S substrates
 * (forward elimination + back substitution, BioFVM_solvers.cpp:186-201).
 */
static void thomas_line(double* rho, size_t base, size_t jump, int n, int S,
                        const double* c1, const double* denom, const double* c)
{
    double* r0 = rho + base * S;
    for (int s = 0; s < S; ++s) r0[s] /= denom[s];
    for (int i = 1; i < n; ++i) {
        double* r = rho + (base + (size_t)i * jump) * S;
        const double* rp = r - jump * S;
        for (int s = 0; s < S; ++s) {
            r[s] += c1[s] * rp[s];
            r[s] /= denom[i * S + s];
        }
    }
    for (int i = n - 2; i >= 0; --i) {
        double* r = rho + (base + (size_t)i * jump) * S;
        const double* rn = r + jump * S;
        for (int s = 0; s < S; ++s) r[s] -= c[i * S + s] * rn[s];
    }
}

/*
 * One LOD step (BioFVM_solvers.cpp:175-273): Dirichlet, x-sweep, Dirichlet,
 * y-sweep, Dirichlet, z-sweep.  Coefficients are recomputed on every call;
 * the reference computes them once per Microenvironment (:87,172), which is
 * the same thing while dt, D, lambda and the grid stay fixed.
 */
void oracle_lod_step(int nx, int ny, int nz, int S, double dx, double dt,
                     const double* D, const double* lambda, double* rho,
                     int n_dir, const int* dir_idx, const double* dir_val,
                     const unsigned char* dir_mask)
{
    double* dnx = malloc(sizeof(double) * (size_t)nx * S);
    double* cx = malloc(sizeof(double) * (size_t)nx * S);
    double* dny = malloc(sizeof(double) * (size_t)ny * S);
    double* cy = malloc(sizeof(double) * (size_t)ny * S);
    double* dnz = malloc(sizeof(double) * (size_t)nz * S);
    double* cz = malloc(sizeof(double) * (size_t)nz * S);
    double* c1 = malloc(sizeof(double) * (size_t)S);
    oracle_thomas_coefficients(nx, S, dx, dt, D, lambda, dnx, cx);
    oracle_thomas_coefficients(ny, S, dx, dt, D, lambda, dny, cy);
    oracle_thomas_coefficients(nz, S, dx, dt, D, lambda, dnz, cz);
    thomas_constant1(S, dx, dt, D, c1);
    const size_t jx = 1, jy = (size_t)nx, jz = (size_t)nx * ny;

    oracle_apply_dirichlet(S, rho, n_dir, dir_idx, dir_val, dir_mask);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            thomas_line(rho, ((size_t)k * ny + j) * nx, jx, nx, S, c1, dnx, cx);

    oracle_apply_dirichlet(S, rho, n_dir, dir_idx, dir_val, dir_mask);
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i)
            thomas_line(rho, (size_t)k * ny * nx + i, jy, ny, S, c1, dny, cy);

    oracle_apply_dirichlet(S, rho, n_dir, dir_idx, dir_val, dir_mask);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            thomas_line(rho, (size_t)j * nx + i, jz, nz, S, c1, dnz, cz);

    free(dnx); free(cx); free(dny); free(cy); free(dnz); free(cz); free(c1);
}

/*
 * One LOD_2D step (diffusion_decay_solver__constant_coefficients_LOD_2D,
 * BioFVM_solvers.cpp:278-427): Dirichlet, x-sweep, Dirichlet, y-sweep, with
 * the decay split in halves.  Only the k = 0 plane is swept (:370, :399).
 */
void oracle_lod2d_step(int nx, int ny, int nz, int S, double dx, double dt,
                       const double* D, const double* lambda, double* rho,
                       int n_dir, const int* dir_idx, const double* dir_val,
                       const unsigned char* dir_mask)
{
    (void)nz;
    double* dnx = malloc(sizeof(double) * (size_t)nx * S);
    double* cx = malloc(sizeof(double) * (size_t)nx * S);
    double* dny = malloc(sizeof(double) * (size_t)ny * S);
    double* cy = malloc(sizeof(double) * (size_t)ny * S);
    double* c1 = malloc(sizeof(double) * (size_t)S);
    thomas_coefficients_split(nx, S, dx, dt, D, lambda, 2, dnx, cx);
    thomas_coefficients_split(ny, S, dx, dt, D, lambda, 2, dny, cy);
    thomas_constant1(S, dx, dt, D, c1);
    oracle_apply_dirichlet(S, rho, n_dir, dir_idx, dir_val, dir_mask);
    for (int j = 0; j < ny; ++j) thomas_line(rho, (size_t)j * nx, 1, nx, S, c1, dnx, cx);
    oracle_apply_dirichlet(S, rho, n_dir, dir_idx, dir_val, dir_mask);
    for (int i = 0; i < nx; ++i) thomas_line(rho, (size_t)i, (size_t)nx, ny, S, c1, dny, cy);
    free(dnx); free(cx); free(dny); free(cy); free(c1);
}

/*
 * Bulk sources/sinks (BioFVM_microenvironment.cpp:595-610):
 *   rho = (rho + dt*(T*S)) / (((U+S)*dt) + 1)
 */
void oracle_bulk_step(size_t nvox, int S, double dt, double* rho,
                      const double* supply, const double* target, const double* uptake)
{
    for (size_t v = 0; v < nvox; ++v) {
        for (int s = 0; s < S; ++s) {
            size_t e = v * S + s;
            double t2 = target[e];
            t2 *= supply[e];               /* temp2 = T*S      :603 */
            double r = rho[e];
            r += dt * t2;                  /* axpy             :604 */
            double t3 = uptake[e];
            t3 += supply[e];               /* U+S              :605 */
            t3 *= dt;                      /* dt*(U+S)         :606 */
            t3 += 1.0;                     /* 1 + dt*(U+S)     :607 */
            r /= t3;                       /*                  :609 */
            rho[e] = r;
        }
    }
}

/*
 * Nearest voxel of a point (BioFVM_mesh.cpp:930-948) for a uniform mesh whose
 * bounding box starts at (x0,y0,z0); returns -1 when the point is outside the
 * closed box, where the reference's assign_position refuses it
 * (BioFVM_basic_agent.cpp:66-71, BioFVM_mesh.cpp:208-217).
 */
int oracle_nearest_voxel(int nx, int ny, int nz, double dx,
                         double x0, double y0, double z0,
                         double x, double y, double z)
{
    if (x < x0 || x > x0 + nx * dx) return -1;
    if (y < y0 || y > y0 + ny * dx) return -1;
    if (z < z0 || z > z0 + nz * dx) return -1;
    int i = (int)floor((x - x0) / dx);
    int j = (int)floor((y - y0) / dx);
    int k = (int)floor((z - z0) / dx);
    if (i >= nx) i = nx - 1;
    if (i < 0) i = 0;
    if (j >= ny) j = ny - 1;
    if (j < 0) j = 0;
    if (k >= nz) k = nz - 1;
    if (k < 0) k = 0;
    return (k * ny + j) * nx + i;
}

/*
 * Per-agent implicit constants (BioFVM_basic_agent.cpp:114-148):
 *   c     = dt*V_cell / V_voxel
 *   temp1 = ((0 + S)*T)*c
 *   temp2 = (1 + c*S) + c*U
 */
void oracle_agent_constants(int n_agents, int S, double dt, double voxel_volume,
                            const double* volume, const double* secretion,
                            const double* saturation, const double* uptake,
                            double* temp1, double* temp2)
{
    for (int a = 0; a < n_agents; ++a) {
        double c = dt * volume[a] / voxel_volume;
        for (int s = 0; s < S; ++s) {
            size_t e = (size_t)a * S + s;
            double t1 = 0.0;
            t1 += secretion[e];
            t1 *= saturation[e];
            t1 *= c;
            double t2 = 1.0;
            t2 += c * secretion[e];
            t2 += c * uptake[e];
            temp1[e] = t1;
            temp2[e] = t2;
        }
    }
}

/*
 * Agent secretion/uptake in list order (BioFVM_basic_agent.cpp:229-238,
 * loop BioFVM_microenvironment.cpp:615-624 run on one thread):
 *   rho[v] = (rho[v] + temp1) / temp2
 */
void oracle_agent_step(int S, double* rho, int n_agents, const int* voxel,
                       const double* temp1, const double* temp2)
{
    for (int a = 0; a < n_agents; ++a) {
        double* r = rho + (size_t)voxel[a] * S;
        for (int s = 0; s < S; ++s) {
            r[s] += temp1[(size_t)a * S + s];
            r[s] /= temp2[(size_t)a * S + s];
        }
    }
}

/*
 * Whole-run driver with the same argument meaning as oracle/_ref's ref_run
 * (plus the optional per-substrate Dirichlet mask): bulk -> cells -> diffusion
 * each step (examples/convergence_test5.cpp:236-238).  The mesh spans
 * [0, n*dx] on each axis, as the reference driver builds it.
 */
int oracle_run(int nx, int ny, int nz, int S, double dx,
               const double* D, const double* lambda, double dt, int steps, int phases,
               double* rho,
               int n_dir, const int* dir_idx, const double* dir_val, const unsigned char* dir_mask,
               const double* bulk_supply, const double* bulk_target, const double* bulk_uptake,
               int n_agents, const double* agent_pos, const double* agent_volume,
               const double* agent_secretion, const double* agent_saturation,
               const double* agent_uptake, int* out_voxel)
{
    size_t nvox = (size_t)nx * ny * nz;
    int* voxel = malloc(sizeof(int) * (size_t)(n_agents > 0 ? n_agents : 1));
    double* t1 = malloc(sizeof(double) * (size_t)(n_agents > 0 ? n_agents : 1) * S);
    double* t2 = malloc(sizeof(double) * (size_t)(n_agents > 0 ? n_agents : 1) * S);
    int rc = 0;
    for (int a = 0; a < n_agents; ++a) {
        voxel[a] = oracle_nearest_voxel(nx, ny, nz, dx, 0.0, 0.0, 0.0, agent_pos[3 * a],
                                        agent_pos[3 * a + 1], agent_pos[3 * a + 2]);
        if (voxel[a] < 0) rc = -2;
        if (out_voxel) out_voxel[a] = voxel[a];
    }
    if (rc == 0 && n_agents > 0) {
        double dV = dx * dx * dx;   /* dx*dy*dz, BioFVM_mesh.cpp:839 */
        oracle_agent_constants(n_agents, S, dt, dV, agent_volume, agent_secretion,
                               agent_saturation, agent_uptake, t1, t2);
    }
    int have_bulk = bulk_supply && bulk_target && bulk_uptake;
    for (int n = 0; rc == 0 && n < steps; ++n) {
        if ((phases & ORACLE_BULK) && have_bulk)
            oracle_bulk_step(nvox, S, dt, rho, bulk_supply, bulk_target, bulk_uptake);
        if ((phases & ORACLE_CELLS) && n_agents > 0)
            oracle_agent_step(S, rho, n_agents, voxel, t1, t2);
        if ((phases & ORACLE_DIFFUSION) && (phases & ORACLE_LOD_2D))
            oracle_lod2d_step(nx, ny, nz, S, dx, dt, D, lambda, rho, n_dir, dir_idx, dir_val, dir_mask);
        else if (phases & ORACLE_DIFFUSION)
            oracle_lod_step(nx, ny, nz, S, dx, dt, D, lambda, rho, n_dir, dir_idx, dir_val, dir_mask);
    }
    free(voxel); free(t1); free(t2);
    return rc;
}

/*
 * Gradient vectors of every voxel (Microenvironment::compute_all_gradient_vectors,
 * BioFVM_microenvironment.cpp:699-779): for each axis, interior voxels only,
 *   g = rho[n + jump]; g -= rho[n - jump]; g /= (2 * d_axis)
 * with 2*d computed as d*2.0 (:701-710).  Components on the two boundary
 * layers of an axis are left untouched; here they are set to 0, the value the
 * reference's reset_all_gradient_vectors gives them (:826-842).
 * out[(n*S + s)*3 + axis].  The reference's two_dx is a function-static set
 * from the first mesh seen (:701-710); here each call uses its own spacing.
 */
void oracle_gradients(int nx, int ny, int nz, int S, double dx, double dy, double dz,
                      const double* rho, double* out)
{
    const size_t nvox = (size_t)nx * ny * nz;
    const int n_axis[3] = {nx, ny, nz};
    const size_t jump[3] = {1, (size_t)nx, (size_t)nx * ny};
    double two_d[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) two_d[a] *= 2.0;
    memset(out, 0, sizeof(double) * nvox * S * 3);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t n = ((size_t)k * ny + j) * nx + i;
                const int coord[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    if (coord[a] < 1 || coord[a] >= n_axis[a] - 1) continue;
                    for (int q = 0; q < S; ++q) {
                        double g = rho[(n + jump[a]) * S + q];
                        g -= rho[(n - jump[a]) * S + q];
                        g /= two_d[a];
                        out[(n * S + q) * 3 + a] = g;
                    }
                }
            }
}
