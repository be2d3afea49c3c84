/*
 * Sanitizer driver for the CPU oracle (TEST INFRASTRUCTURE). Built by
 * `make -C oracle sanitize` together with dp_oracle.c and nep_oracle.c under
 * AddressSanitizer and UndefinedBehaviorSanitizer (host code only, no OpenMP)
 * and run by tests/test_oracle_sanitize.py: every oracle routine runs once on
 * small frames whose buffers are sized exactly as the ctypes layer sizes them,
 * so an out-of-bounds access, a use of uninitialised heap through a bad index,
 * or undefined arithmetic aborts the run. It also checks two invariants of the
 * results (the forces of a periodic frame sum to zero; every neighbor of a
 * symmetric list names its centre back) so that a silently wrong run fails.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_dp_find_neighbor(const double *lattice, double *frac, const int32_t *type_map, int ntypes, int natoms,
                            int m, double rc_m, const double *rc_type, int32_t *list, double *dR);
void oracle_dp_image_dr(const int32_t *list, const double *dR, long nslots, double *image_dr);
void oracle_dp_env_mat(const double *image_dr, const int32_t *type_map, const double *davg, const double *dstd,
                       int natoms, int S, double rmin, double rmax, double *Ri, double *Ri_d);
void oracle_dp_calc_stat(const int32_t *list, const double *dR, const int32_t *type_map, int natoms, int S,
                         int ntypes, double rmin, double rmax, double *davg, double *dstd);
int oracle_dp_energy_grad(const double *Ri, const int32_t *type_map, const int32_t *type_present, int natoms,
                          int ntypes, int m, int M2, const int32_t *emb_desc, const double *emb_params,
                          const int32_t *fit_desc, const double *fit_params, double *Ei, double *dE);
void oracle_dp_force_virial(const int32_t *list, const double *dE, const double *Ri_d, const double *image_dr,
                            int natoms, int nghost, int S, double *force, double *virial);
void oracle_dp_force_op(const int32_t *nlist, const double *dE, const double *Ri_d, const double *F_center, int B,
                        int N, int S, int nghost, double *force);
void oracle_dp_virial_op(const int32_t *nlist, const double *dE, const double *Rij, const double *Ri_d, int B,
                         int N, int S, int nghost, double *virial, double *atom_virial);
void oracle_dp_force_grad_op(const int32_t *nlist, const double *Ri_d, const double *gF, int B, int N, int S,
                             int nghost, double *g);
void oracle_dp_virial_grad_op(const int32_t *nlist, const double *Rij, const double *Ri_d, const double *gV, int B,
                              int N, int S, double *g);
double oracle_nep_expand_box(const double *lat, double cutoff, int64_t *num_cell, double *box);
int oracle_nep_neighbor(int64_t nframes, const int64_t *num_atoms, const double *box, const double *box_orig,
                        const int64_t *num_cell, const double *pos, double rc2b, double rc3b, int64_t maxr,
                        int64_t maxa, int64_t ntot, int64_t *NN_r, int64_t *NL_r, int64_t *NN_a, int64_t *NL_a,
                        double *Ri_r, double *Ri_a);
void oracle_nep_maxneigh(int64_t nframes, const int64_t *num_atoms, const double *box, const double *box_orig,
                         const int64_t *num_cell, const double *pos, double rc2b, double rc3b, int ntypes,
                         const int64_t *type_map, int with_type, int64_t *NN_r, int64_t *NN_a);

static uint64_t rng_state = 88172645463325252ull;
static double urand(void) /* xorshift64, uniform in [0, 1) */
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static void *xcalloc(size_t n, size_t size)
{
    void *p = calloc(n ? n : 1, size);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static int fail(const char *what, double v)
{
    fprintf(stderr, "san_driver: %s (%g)\n", what, v);
    return 1;
}

/* fcc cell of side a, reps^3 cells, types alternating by sublattice when ntypes = 2 */
static int fcc(int reps, double a, int ntypes, double *L, double *frac, int32_t *types)
{
    static const double basis[4][3] = {{0, 0, 0}, {0, .5, .5}, {.5, 0, .5}, {.5, .5, 0}};
    int n = 0;
    memset(L, 0, 9 * sizeof(double));
    L[0] = L[4] = L[8] = a * reps;
    for (int t = 0; t < ntypes; t++)
        for (int x = 0; x < reps; x++)
            for (int y = 0; y < reps; y++)
                for (int z = 0; z < reps; z++)
                    for (int b = 0; b < 4; b++) {
                        if (ntypes == 2 && (b % 2) != t) continue;
                        if (ntypes == 1 && t > 0) continue;
                        frac[3 * n + 0] = (x + basis[b][0] + 0.02 * (urand() - 0.5)) / reps;
                        frac[3 * n + 1] = (y + basis[b][1] + 0.02 * (urand() - 0.5)) / reps;
                        frac[3 * n + 2] = (z + basis[b][2] + 0.02 * (urand() - 0.5)) / reps;
                        types[n++] = t;
                    }
    return n;
}

static void net_params(int nl, const int *sz, int kind, int32_t *desc, double **params)
{
    desc[0] = nl;
    for (int l = 0; l <= nl; l++) desc[1 + l] = sz[l];
    desc[2 + nl] = 1; /* bias */
    desc[3 + nl] = 1; /* resnet_dt where the layer rules give one */
    size_t cnt = 0;
    for (int l = 0; l < nl; l++) {
        cnt += (size_t)sz[l] * sz[l + 1] + sz[l + 1];
        const int rdt = kind == 0 ? (sz[l + 1] == sz[l] || sz[l + 1] == 2 * sz[l]) : (l > 0 && l < nl - 1);
        if (rdt) cnt += sz[l + 1];
    }
    double *p = xcalloc(cnt, sizeof(double));
    for (size_t k = 0; k < cnt; k++) p[k] = 0.3 * (urand() - 0.5);
    *params = p;
}

static int run_dp(int reps, int ntypes)
{
    const int m = 100, M2 = 4, E = 8;
    const double rc = 6.0, rmin = 5.5;
    double L[9];
    const int nmax = 4 * reps * reps * reps;
    double *frac = xcalloc(3 * (size_t)nmax, sizeof(double));
    int32_t *types = xcalloc(nmax, sizeof(int32_t));
    const int N = fcc(reps, 3.615, ntypes, L, frac, types);
    const int S = ntypes * m;
    double rct[2] = {rc, rc};
    int32_t *list = xcalloc((size_t)N * S, sizeof(int32_t));
    double *dR = xcalloc((size_t)N * S * 3, sizeof(double));
    if (oracle_dp_find_neighbor(L, frac, types, ntypes, N, m, rc, rct, list, dR)) return fail("overflow", N);
    /* every listed neighbor names its centre back */
    for (int i = 0; i < N; i++)
        for (int s = 0; s < S; s++) {
            const int j = list[(size_t)i * S + s] - 1;
            if (j < 0) continue;
            int back = 0;
            for (int u = 0; u < S && !back; u++) back = list[(size_t)j * S + u] - 1 == i;
            if (!back) return fail("asymmetric list", i);
        }
    double *idr = xcalloc((size_t)N * S * 4, sizeof(double));
    oracle_dp_image_dr(list, dR, (long)N * S, idr);
    double davg[8], dstd[8];
    oracle_dp_calc_stat(list, dR, types, N, S, ntypes, rmin, rc, davg, dstd);
    double *davg_full = xcalloc((size_t)ntypes * S * 4, sizeof(double));
    double *dstd_full = xcalloc((size_t)ntypes * S * 4, sizeof(double));
    for (int t = 0; t < ntypes; t++)
        for (int s = 0; s < S; s++)
            for (int w = 0; w < 4; w++) {
                davg_full[((size_t)t * S + s) * 4 + w] = davg[4 * t + w];
                dstd_full[((size_t)t * S + s) * 4 + w] = dstd[4 * t + w];
            }
    double *Ri = xcalloc((size_t)N * S * 4, sizeof(double));
    double *Ri_d = xcalloc((size_t)N * S * 12, sizeof(double));
    oracle_dp_env_mat(idr, types, davg_full, dstd_full, N, S, rmin, rc, Ri, Ri_d);
    const int esz[4] = {1, E, E, E}, fsz[5] = {E * M2, 10, 10, 10, 1};
    int32_t edesc[16], fdesc[16];
    double *ep_one, *fp_one;
    net_params(3, esz, 0, edesc, &ep_one);
    net_params(4, fsz, 1, fdesc, &fp_one);
    /* every pair and type gets the same net; block sizes as the oracle binds them: bias always,
     * resnet_dt by the layer rules */
    size_t eb = 0, fb = 0;
    for (int l = 0; l < 3; l++)
        eb += (size_t)esz[l] * esz[l + 1] + esz[l + 1] +
              ((esz[l + 1] == esz[l] || esz[l + 1] == 2 * esz[l]) ? esz[l + 1] : 0);
    for (int l = 0; l < 4; l++) fb += (size_t)fsz[l] * fsz[l + 1] + fsz[l + 1] + ((l > 0 && l < 3) ? fsz[l + 1] : 0);
    double *ep = xcalloc(eb * ntypes * ntypes, sizeof(double));
    double *fp = xcalloc(fb * ntypes, sizeof(double));
    for (int p = 0; p < ntypes * ntypes; p++) memcpy(ep + p * eb, ep_one, eb * sizeof(double));
    for (int t = 0; t < ntypes; t++) memcpy(fp + t * fb, fp_one, fb * sizeof(double));
    int32_t present[2] = {1, 1};
    double *Ei = xcalloc(N, sizeof(double));
    double *dE = xcalloc((size_t)N * S * 4, sizeof(double));
    if (oracle_dp_energy_grad(Ri, types, present, N, ntypes, m, M2, edesc, ep, fdesc, fp, Ei, dE))
        return fail("energy_grad", 0);
    double *F = xcalloc((size_t)N * 3, sizeof(double));
    double V[9];
    oracle_dp_force_virial(list, dE, Ri_d, idr, N, 0, S, F, V);
    double fs[3] = {0, 0, 0}, fmax = 0;
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            fs[c] += F[3 * i + c];
            if (fabs(F[3 * i + c]) > fmax) fmax = fabs(F[3 * i + c]);
        }
    for (int c = 0; c < 3; c++)
        if (fabs(fs[c]) > 1e-10 * (fmax > 1 ? fmax : 1) * N) return fail("forces do not sum to zero", fs[c]);
    /* the operator contract on the same list (0-based, -1 empty), one ghost row */
    const int ng = 1;
    int32_t *nl0 = xcalloc((size_t)N * S, sizeof(int32_t));
    double *Rij = xcalloc((size_t)N * S * 3, sizeof(double));
    for (size_t k = 0; k < (size_t)N * S; k++) {
        nl0[k] = list[k] - 1;
        for (int c = 0; c < 3; c++) Rij[3 * k + c] = dR[3 * k + c];
    }
    double *Fout = xcalloc((size_t)(N + ng) * 3, sizeof(double));
    oracle_dp_force_op(nl0, dE, Ri_d, F, 1, N, S, ng, Fout);
    double vir[9], *av = xcalloc((size_t)(N + ng) * 9, sizeof(double));
    oracle_dp_virial_op(nl0, dE, Rij, Ri_d, 1, N, S, ng, vir, av);
    double *gF = xcalloc((size_t)(N + ng) * 3, sizeof(double)), gV[9];
    for (int k = 0; k < (N + ng) * 3; k++) gF[k] = urand() - 0.5;
    for (int k = 0; k < 9; k++) gV[k] = urand() - 0.5;
    double *g = xcalloc((size_t)N * S * 4, sizeof(double));
    oracle_dp_force_grad_op(nl0, Ri_d, gF, 1, N, S, ng, g);
    oracle_dp_virial_grad_op(nl0, Rij, Ri_d, gV, 1, N, S, g);
    free(frac); free(types); free(list); free(dR); free(idr); free(davg_full); free(dstd_full); free(Ri);
    free(Ri_d); free(ep_one); free(fp_one); free(ep); free(fp); free(Ei); free(dE); free(F); free(nl0);
    free(Rij); free(Fout); free(av); free(gF); free(g);
    printf("dp: %d atoms, %d types, sum F = (%.1e, %.1e, %.1e)\n", N, ntypes, fs[0], fs[1], fs[2]);
    return 0;
}

static int run_nep(void)
{
    /* two frames: a small cell (replicated by expand_box) and a larger one */
    const int reps[2] = {1, 2};
    int64_t num[2], nc[6];
    double box[36], orig[18];
    double *pos = xcalloc(3 * 4 * 8 * 2, sizeof(double));
    int64_t *tm = xcalloc(4 * 8 * 2, sizeof(int64_t));
    int64_t ntot = 0;
    for (int f = 0; f < 2; f++) {
        double L[9], frac[3 * 32];
        int32_t types[32];
        const int n = fcc(reps[f], 4.36, 1, L, frac, types);
        num[f] = n;
        /* column-major lattice as expand_box takes it (the vectors are the columns) */
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) orig[9 * f + 3 * r + c] = L[3 * c + r];
        oracle_nep_expand_box(orig + 9 * f, 6.0, nc + 3 * f, box + 18 * f);
        for (int i = 0; i < n; i++) {
            for (int c = 0; c < 3; c++)
                pos[3 * (ntot + i) + c] = frac[3 * i] * L[c] + frac[3 * i + 1] * L[3 + c] + frac[3 * i + 2] * L[6 + c];
            tm[ntot + i] = i % 2;
        }
        ntot += n;
    }
    int64_t *NNr = xcalloc(ntot * 2, sizeof(int64_t)), *NNa = xcalloc(ntot * 2, sizeof(int64_t));
    oracle_nep_maxneigh(2, num, box, orig, nc, pos, 6.0, 4.0, 2, tm, 1, NNr, NNa);
    oracle_nep_maxneigh(2, num, box, orig, nc, pos, 6.0, 4.0, 2, tm, 0, NNr, NNa);
    int64_t maxr = 1, maxa = 1;
    for (int64_t i = 0; i < ntot; i++) {
        if (NNr[i] > maxr) maxr = NNr[i];
        if (NNa[i] > maxa) maxa = NNa[i];
    }
    int64_t *NLr = xcalloc(ntot * maxr, sizeof(int64_t)), *NLa = xcalloc(ntot * maxa, sizeof(int64_t));
    double *Rr = xcalloc(ntot * maxr * 4, sizeof(double)), *Ra = xcalloc(ntot * maxa * 4, sizeof(double));
    int64_t *N2r = xcalloc(ntot, sizeof(int64_t)), *N2a = xcalloc(ntot, sizeof(int64_t));
    if (oracle_nep_neighbor(2, num, box, orig, nc, pos, 6.0, 4.0, maxr, maxa, ntot, N2r, NLr, N2a, NLa, Rr, Ra))
        return fail("nep neighbor overflow", 0);
    for (int64_t i = 0; i < ntot; i++)
        if (N2r[i] != NNr[i] || N2a[i] != NNa[i]) return fail("nep counts differ from maxneigh", (double)i);
    printf("nep: %lld atoms, max neighbors %lld / %lld\n", (long long)ntot, (long long)maxr, (long long)maxa);
    free(pos); free(tm); free(NNr); free(NNa); free(NLr); free(NLa); free(Rr); free(Ra); free(N2r); free(N2a);
    return 0;
}

int main(void)
{
    int rc = run_dp(2, 1);   /* 32 atoms: a box below 2 rc, every image enumerated */
    rc |= run_dp(4, 2);      /* 256 atoms, two types: the nearest-image rule */
    rc |= run_nep();
    if (!rc) printf("san_driver: ok\n");
    return rc;
}
