/*
 * FP64 CPU oracle for the Deep Potential se_e2_a descriptor-and-force path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP kernels in
 * matpl_amd/csrc are compared against; it is never linked into, loaded by or
 * called from the product library. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it.
 *
 * Each routine restates one piece of the reference algorithm and cites the
 * reference file:line it follows:
 *   oracle_dp_find_neighbor  src/pre_data/gen_feature/NeighConst.f90:14-149
 *   oracle_dp_image_dr       src/pre_data/dpuni_data_loader.py:631-649,652-716
 *   oracle_dp_env_mat        src/model/dp_dp.py:217-302
 *   oracle_dp_calc_stat      src/pre_data/dpuni_data_loader.py:422-560
 *   oracle_dp_energy_grad    src/model/dp_dp.py:304-429, src/model/dp_embedding.py:40-160
 *   oracle_dp_force_virial   src/model/dp_dp.py:436-459 (CPU branch)
 *   oracle_dp_force_op       src/op/kernel/calculateForce.cu:49-93 (+ CalcOps.cpp:6-23)
 *   oracle_dp_virial_op      src/op/kernel/calculateVirial.cu:18-120 (+ CalcOps.cpp:170-186)
 *   oracle_dp_force_grad_op  src/op/kernel/calculateForceGrad.cu:9-90
 *   oracle_dp_virial_grad_op src/op/kernel/calculateVirialGrad.cu:15-92
 *
 * Compiled with -ffp-contract=off so that every FP64 expression rounds the
 * way it is written; the neighbor displacement in the HIP kernel is written
 * in the same order so that neighbor lists and dR match bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_LAYERS 8
#define ORACLE_MAX_WIDTH 1024

/* ------------------------------------------------------------------------ */
/* D1: neighbor list, NeighConst.f90:14-149                                  */
/* ------------------------------------------------------------------------ */
/*
 * lattice : [3][3] row i = lattice vector i (Angstrom), as lattice(n,i,:)
 * frac    : [natoms][3] fractional coordinates; wrapped in place with
 *           mod(x+10,1) exactly as NeighConst.f90:46-50 does
 * type_map: [natoms] 0-based type of each atom
 * rc_type : [ntypes] per-center-type cutoff (Rc_type)
 * list    : [natoms][ntypes][m] out, 1-based neighbor index, 0 = empty
 * dR      : [natoms][ntypes][m][3] out, r_j - r_i (Cartesian)
 * returns 0, or 1 + (index of the first atom whose bucket overflowed m)
 */
int oracle_dp_find_neighbor(const double *lattice, double *frac, const int32_t *type_map,
                            int ntypes, int natoms, int m, double rc_m,
                            const double *rc_type, int32_t *list, double *dR)
{
    int nperiod[3];
    int small_box = 0;
    for (int i = 0; i < 3; i++) {
        const double *a = lattice + 3 * i;
        double d = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        nperiod[i] = (int)(rc_m / d) + 1;
        if (d < 2.0 * rc_m) small_box = 1;
    }
    for (int i = 0; i < natoms * 3; i++) frac[i] = fmod(frac[i] + 10.0, 1.0);

    const double rc2 = rc_m * rc_m;
    const size_t S = (size_t)ntypes * m;
    memset(list, 0, sizeof(int32_t) * natoms * S);
    memset(dR, 0, sizeof(double) * natoms * S * 3);
    int err = 0;

#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < natoms; i++) {
        int cnt[64], cntM[64];
        for (int t = 0; t < ntypes; t++) cnt[t] = cntM[t] = 0;
        const double rci = rc_type[type_map[i]];
        const double rci2 = rci * rci;
        int32_t *li = list + (size_t)i * S;
        double *di = dR + (size_t)i * S * 3;
        int bad = 0;
        for (int j = 0; j < natoms && !bad; j++) {
            const int tj = type_map[j];
            int lo1 = small_box ? -nperiod[0] : 0, hi1 = small_box ? nperiod[0] : 0;
            int lo2 = small_box ? -nperiod[1] : 0, hi2 = small_box ? nperiod[1] : 0;
            int lo3 = small_box ? -nperiod[2] : 0, hi3 = small_box ? nperiod[2] : 0;
            for (int i1 = lo1; i1 <= hi1; i1++)
            for (int i2 = lo2; i2 <= hi2; i2++)
            for (int i3 = lo3; i3 <= hi3; i3++) {
                double dx1 = frac[3 * j + 0] - frac[3 * i + 0];
                double dx2 = frac[3 * j + 1] - frac[3 * i + 1];
                double dx3 = frac[3 * j + 2] - frac[3 * i + 2];
                if (small_box) {
                    dx1 = dx1 + (double)i1;
                    dx2 = dx2 + (double)i2;
                    dx3 = dx3 + (double)i3;
                } else {
                    /* NeighConst.f90:98-108: nearest of dx, dx+1, dx-1 */
                    if (fabs(dx1 + 1.0) < fabs(dx1)) dx1 = dx1 + 1.0;
                    if (fabs(dx1 - 1.0) < fabs(dx1)) dx1 = dx1 - 1.0;
                    if (fabs(dx2 + 1.0) < fabs(dx2)) dx2 = dx2 + 1.0;
                    if (fabs(dx2 - 1.0) < fabs(dx2)) dx2 = dx2 - 1.0;
                    if (fabs(dx3 + 1.0) < fabs(dx3)) dx3 = dx3 + 1.0;
                    if (fabs(dx3 - 1.0) < fabs(dx3)) dx3 = dx3 - 1.0;
                }
                double dx = lattice[0] * dx1 + lattice[3] * dx2 + lattice[6] * dx3;
                double dy = lattice[1] * dx1 + lattice[4] * dx2 + lattice[7] * dx3;
                double dz = lattice[2] * dx1 + lattice[5] * dx2 + lattice[8] * dx3;
                double dd = dx * dx + dy * dy + dz * dz;
                if (dd < rc2 && dd > 1e-8) {
                    cntM[tj]++;
                    if (cntM[tj] > m) { bad = 1; break; }
                    if (dd < rci2) {
                        int k = cnt[tj]++;
                        li[tj * m + k] = j + 1;
                        di[3 * (tj * m + k) + 0] = dx;
                        di[3 * (tj * m + k) + 1] = dy;
                        di[3 * (tj * m + k) + 2] = dz;
                    }
                }
            }
        }
        if (bad) {
#pragma omp critical
            { if (err == 0 || i + 1 < err) err = i + 1; }
        }
    }
    return err;
}

/* ImageDR = [r, dx, dy, dz]; r only where list > 0 (dpuni_data_loader.py:669-673) */
void oracle_dp_image_dr(const int32_t *list, const double *dR, long nslots, double *image_dr)
{
    for (long s = 0; s < nslots; s++) {
        const double *d = dR + 3 * s;
        double *o = image_dr + 4 * s;
        o[1] = d[0]; o[2] = d[1]; o[3] = d[2];
        o[0] = list[s] > 0 ? sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) : 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* D3: environment matrix Ri and its derivative, dp_dp.py:217-302           */
/* ------------------------------------------------------------------------ */
static void smooth_s(double r, double rmin, double rmax, double *s, double *ds)
{
    /* dp_dp.py:244-249: s = 1/r (0<r<rmin); (1/r)(u^3(-6u^2+15u-10)+1) (rmin<r<rmax) */
    if (r > 0.0 && r < rmin) {
        *s = 1.0 / r;
        *ds = -1.0 / (r * r);
    } else if (r > rmin && r < rmax) {
        double u = (r - rmin) / (rmax - rmin);
        double vv = u * u * u * (-6.0 * u * u + 15.0 * u - 10.0) + 1.0;
        double dvv = -30.0 * u * u * (u - 1.0) * (u - 1.0) / (rmax - rmin);
        *s = vv / r;
        *ds = dvv / r - vv / (r * r);
    } else {
        *s = 0.0;
        *ds = 0.0;
    }
}

/*
 * image_dr: [N][S][4]; type_map [N]; davg/dstd: [T][S][4] (the reference
 * layout, dp_dp.py:293-296); out Ri [N][S][4] normalized, Ri_d [N][S][4][3]
 * = -dfeat/dstd (dp_dp.py:297-299).
 */
void oracle_dp_env_mat(const double *image_dr, const int32_t *type_map, const double *davg,
                       const double *dstd, int natoms, int S, double rmin, double rmax,
                       double *Ri, double *Ri_d)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < natoms; i++) {
        const int ti = type_map[i];
        for (int k = 0; k < S; k++) {
            const double *d = image_dr + ((size_t)i * S + k) * 4;
            const double r = d[0];
            double s, ds;
            smooth_s(r, rmin, rmax, &s, &ds);
            double ri[4] = {s, 0.0, 0.0, 0.0};
            double df[4][3];
            memset(df, 0, sizeof(df));
            if (fabs(r) > 1e-5) {
                const double x[3] = {d[1], d[2], d[3]};
                ri[1] = s * x[0] / r;
                ri[2] = s * x[1] / r;
                ri[3] = s * x[2] / r;
                const double common = (ds - s / r) / (r * r);
                for (int a = 0; a < 3; a++) {
                    df[0][a] = ds * x[a] / r;
                    for (int b = 0; b < 3; b++)
                        df[b + 1][a] = common * x[b] * x[a] + (a == b ? s / r : 0.0);
                }
            }
            const double *av = davg + ((size_t)ti * S + k) * 4;
            const double *sd = dstd + ((size_t)ti * S + k) * 4;
            double *ro = Ri + ((size_t)i * S + k) * 4;
            double *rdo = Ri_d + ((size_t)i * S + k) * 12;
            for (int w = 0; w < 4; w++) {
                ro[w] = (ri[w] - av[w]) / sd[w];
                for (int a = 0; a < 3; a++) rdo[w * 3 + a] = -df[w][a] / sd[w];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* D4: davg / dstd statistics, dpuni_data_loader.py:422-560                  */
/* ------------------------------------------------------------------------ */
static double compute_std(double sum2, double sum, double n)
{
    if (n == 0.0) return 1e-2;
    double val = sqrt(sum2 / n - (sum / n) * (sum / n));
    if (fabs(val) < 1e-2) val = 1e-2;
    return val;
}

/*
 * Statistics of the un-normalized Ri over every slot (empty slots included,
 * as the reference averages over natoms_of_type * ntypes * m rows).
 * list/dR: [N][S][3]; out davg/dstd [T][4].
 */
void oracle_dp_calc_stat(const int32_t *list, const double *dR, const int32_t *type_map,
                         int natoms, int S, int ntypes, double rmin, double rmax,
                         double *davg, double *dstd)
{
    for (int t = 0; t < ntypes; t++) {
        double sum[4] = {0, 0, 0, 0}, sum2[4] = {0, 0, 0, 0};
        double n = 0.0;
        for (int i = 0; i < natoms; i++) {
            if (type_map[i] != t) continue;
            for (int k = 0; k < S; k++) {
                n += 1.0;
                const size_t s = (size_t)i * S + k;
                if (list[s] <= 0) continue;
                const double *d = dR + 3 * s;
                const double dr2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double r = sqrt(dr2);
                const double x = dr2 / r; /* nr in the reference */
                double vv;
                if (x < rmin) vv = 1.0;
                else if (x < rmax) {
                    double u = (x - rmin) / (rmax - rmin);
                    vv = u * u * u * (-6.0 * u * u + 15.0 * u - 10.0) + 1.0;
                } else vv = 0.0;
                double ri[4] = {vv / x, vv * d[0] / dr2, vv * d[1] / dr2, vv * d[2] / dr2};
                for (int w = 0; w < 4; w++) { sum[w] += ri[w]; sum2[w] += ri[w] * ri[w]; }
            }
        }
        const double sa = (sum[1] + sum[2] + sum[3]) / 3.0;
        const double sa2 = (sum2[1] + sum2[2] + sum2[3]) / 3.0;
        davg[4 * t + 0] = sum[0] / (n + 1e-15);
        davg[4 * t + 1] = davg[4 * t + 2] = davg[4 * t + 3] = 0.0;
        dstd[4 * t + 0] = compute_std(sum2[0], sum[0], n);
        dstd[4 * t + 1] = dstd[4 * t + 2] = dstd[4 * t + 3] = compute_std(sa2, sa, n);
    }
}

/* ------------------------------------------------------------------------ */
/* D5: nets (dp_embedding.py:40-160) and their backward                      */
/* ------------------------------------------------------------------------ */
/*
 * A net is described by an int array
 *   desc = {nl, sz[0..nl], has_bias, has_rdt}
 * and a flat parameter block: for each layer l, W[sz[l]][sz[l+1]] (row-major,
 * x @ W), then b[sz[l+1]] if has_bias, then rdt[sz[l+1]] if the layer owns a
 * resnet_dt vector. Layer rules:
 *   kind 0 (embedding, dp_embedding.py:73-94): out==in -> x = h(*rdt) + x;
 *          out==2in -> x = [x,x] + h(*rdt); else x = h.
 *   kind 1 (fitting, dp_embedding.py:140-163): layer 0 -> x = h; hidden layers
 *          l>0 -> x = h*rdt + x (rdt present for l>0 when has_rdt) or h + x;
 *          the last layer is linear.
 */
typedef struct {
    int kind, nl, sz[ORACLE_MAX_LAYERS + 1], has_bias, has_rdt;
    const double *W[ORACLE_MAX_LAYERS], *b[ORACLE_MAX_LAYERS], *rdt[ORACLE_MAX_LAYERS];
} net_t;

static size_t net_bind(net_t *n, int kind, const int32_t *desc, const double *p)
{
    n->kind = kind;
    n->nl = desc[0];
    for (int l = 0; l <= n->nl; l++) n->sz[l] = desc[1 + l];
    n->has_bias = desc[2 + n->nl];
    n->has_rdt = desc[3 + n->nl];
    const double *q = p;
    for (int l = 0; l < n->nl; l++) {
        const int in = n->sz[l], out = n->sz[l + 1];
        n->W[l] = q; q += (size_t)in * out;
        n->b[l] = NULL; n->rdt[l] = NULL;
        if (n->has_bias) { n->b[l] = q; q += out; }
        int owns_rdt;
        if (kind == 0) owns_rdt = n->has_rdt && (out == in || out == 2 * in);
        else owns_rdt = n->has_rdt && l > 0 && l < n->nl - 1;
        if (owns_rdt) { n->rdt[l] = q; q += out; }
    }
    return (size_t)(q - p);
}

/* Forward; acts[l] keeps the input of layer l (x_l) and tanh outputs t_l. */
typedef struct {
    double x[ORACLE_MAX_LAYERS + 1][ORACLE_MAX_WIDTH];
    double t[ORACLE_MAX_LAYERS][ORACLE_MAX_WIDTH];
} net_ws_t;

static void net_forward(const net_t *n, const double *in, net_ws_t *ws, double *out)
{
    memcpy(ws->x[0], in, sizeof(double) * n->sz[0]);
    for (int l = 0; l < n->nl; l++) {
        const int ni = n->sz[l], no = n->sz[l + 1];
        const double *x = ws->x[l];
        double *y = ws->x[l + 1];
        const int last = (n->kind == 1 && l == n->nl - 1);
        for (int o = 0; o < no; o++) {
            double z = 0.0;
            for (int k = 0; k < ni; k++) z += x[k] * n->W[l][(size_t)k * no + o];
            if (n->b[l]) z += n->b[l][o];
            if (last) { y[o] = z; continue; }
            const double h = tanh(z);
            ws->t[l][o] = h;
            const double hs = n->rdt[l] ? h * n->rdt[l][o] : h;
            if (n->kind == 0) {
                if (no == ni) y[o] = hs + x[o];
                else if (no == 2 * ni) y[o] = x[o % ni] + hs;
                else y[o] = h;
            } else {
                y[o] = (l == 0) ? h : hs + x[o];
            }
        }
    }
    memcpy(out, ws->x[n->nl], sizeof(double) * n->sz[n->nl]);
}

/* Backward of net_forward: gout [sz[nl]] -> gin [sz[0]]. */
static void net_backward(const net_t *n, const net_ws_t *ws, const double *gout, double *gin)
{
    double g[ORACLE_MAX_WIDTH], gprev[ORACLE_MAX_WIDTH], gz[ORACLE_MAX_WIDTH];
    memcpy(g, gout, sizeof(double) * n->sz[n->nl]);
    for (int l = n->nl - 1; l >= 0; l--) {
        const int ni = n->sz[l], no = n->sz[l + 1];
        const int last = (n->kind == 1 && l == n->nl - 1);
        for (int k = 0; k < ni; k++) gprev[k] = 0.0;
        for (int o = 0; o < no; o++) {
            if (last) { gz[o] = g[o]; continue; }
            const double h = ws->t[l][o];
            const double dh = 1.0 - h * h;
            double gh;
            if (n->kind == 0) {
                if (no == ni) { gh = n->rdt[l] ? g[o] * n->rdt[l][o] : g[o]; gprev[o] += g[o]; }
                else if (no == 2 * ni) { gh = n->rdt[l] ? g[o] * n->rdt[l][o] : g[o]; gprev[o % ni] += g[o]; }
                else gh = g[o];
            } else {
                if (l == 0) gh = g[o];
                else { gh = n->rdt[l] ? g[o] * n->rdt[l][o] : g[o]; gprev[o] += g[o]; }
            }
            gz[o] = gh * dh;
        }
        for (int k = 0; k < ni; k++) {
            double acc = 0.0;
            for (int o = 0; o < no; o++) acc += n->W[l][(size_t)k * no + o] * gz[o];
            gprev[k] += acc;
        }
        memcpy(g, gprev, sizeof(double) * ni);
    }
    memcpy(gin, g, sizeof(double) * n->sz[0]);
}

/*
 * D5+D6: per-atom energy and dE/dRi.
 * Ri [N][S][4] normalized; type_map [N]; type_present [T] (1 if the type is in
 * the frame: dp_dp.py:150-165 builds emb_list / type_nums from it);
 * emb_desc/emb_params: T*T embedding nets, stored in order ti*T+tj;
 * fit_desc/fit_params: T fitting nets; M2 axis neurons; m = maxNeighborNum.
 * Out: Ei [N] (in atom order), dE [N][S][4] = dEtot/dRi.
 * Follows dp_dp.py:344-415: A = sum_t Ri_t^T G_t / (m*type_nums) over every
 * slot of the present neighbor types (padding included), DR = A^T A[:, :M2].
 */
int oracle_dp_energy_grad(const double *Ri, const int32_t *type_map, const int32_t *type_present,
                          int natoms, int ntypes, int m, int M2,
                          const int32_t *emb_desc, const double *emb_params,
                          const int32_t *fit_desc, const double *fit_params,
                          double *Ei, double *dE)
{
    const int S = ntypes * m;
    net_t emb[64], fit[8];
    if (ntypes * ntypes > 64 || ntypes > 8) return -1;
    size_t off = 0;
    for (int p = 0; p < ntypes * ntypes; p++) off += net_bind(&emb[p], 0, emb_desc, emb_params + off);
    off = 0;
    for (int t = 0; t < ntypes; t++) off += net_bind(&fit[t], 1, fit_desc, fit_params + off);
    const int E = emb[0].sz[emb[0].nl];
    if (fit[0].sz[0] != E * M2 || E > ORACLE_MAX_WIDTH) return -2;
    int type_nums = 0;
    for (int t = 0; t < ntypes; t++) type_nums += type_present[t] ? 1 : 0;
    const double scale = 1.0 / ((double)m * type_nums);

#pragma omp parallel
    {
        net_ws_t *ws = (net_ws_t *)malloc(sizeof(net_ws_t));
        double *A = (double *)malloc(sizeof(double) * 4 * E);
        double *dA = (double *)malloc(sizeof(double) * 4 * E);
        double *DR = (double *)malloc(sizeof(double) * E * M2);
        double *dDR = (double *)malloc(sizeof(double) * E * M2);
        double G[ORACLE_MAX_WIDTH], dG[ORACLE_MAX_WIDTH];
#pragma omp for schedule(dynamic, 8)
        for (int i = 0; i < natoms; i++) {
            const int ti = type_map[i];
            const double *ri = Ri + (size_t)i * S * 4;
            double *dei = dE + (size_t)i * S * 4;
            memset(dei, 0, sizeof(double) * S * 4);
            memset(A, 0, sizeof(double) * 4 * E);
            for (int tj = 0; tj < ntypes; tj++) {
                if (!type_present[tj]) continue;
                const net_t *en = &emb[ti * ntypes + tj];
                for (int k = 0; k < m; k++) {
                    const double *r4 = ri + (size_t)(tj * m + k) * 4;
                    net_forward(en, r4, ws, G);
                    for (int w = 0; w < 4; w++)
                        for (int c = 0; c < E; c++) A[w * E + c] += r4[w] * G[c];
                }
            }
            for (int x = 0; x < 4 * E; x++) A[x] *= scale;
            for (int p = 0; p < E; p++)
                for (int q = 0; q < M2; q++) {
                    double acc = 0.0;
                    for (int w = 0; w < 4; w++) acc += A[w * E + p] * A[w * E + q];
                    DR[p * M2 + q] = acc;
                }
            double e;
            net_forward(&fit[ti], DR, ws, &e);
            Ei[i] = e;
            double one = 1.0;
            net_backward(&fit[ti], ws, &one, dDR);
            for (int w = 0; w < 4; w++)
                for (int c = 0; c < E; c++) {
                    double acc = 0.0;
                    for (int q = 0; q < M2; q++) acc += dDR[c * M2 + q] * A[w * E + q];
                    if (c < M2)
                        for (int p = 0; p < E; p++) acc += dDR[p * M2 + c] * A[w * E + p];
                    dA[w * E + c] = acc * scale;
                }
            for (int tj = 0; tj < ntypes; tj++) {
                if (!type_present[tj]) continue;
                const net_t *en = &emb[ti * ntypes + tj];
                for (int k = 0; k < m; k++) {
                    const double *r4 = ri + (size_t)(tj * m + k) * 4;
                    double *d4 = dei + (size_t)(tj * m + k) * 4;
                    net_forward(en, r4, ws, G);
                    for (int c = 0; c < E; c++) {
                        double acc = 0.0;
                        for (int w = 0; w < 4; w++) acc += dA[w * E + c] * r4[w];
                        dG[c] = acc;
                    }
                    for (int w = 0; w < 4; w++) {
                        double acc = 0.0;
                        for (int c = 0; c < E; c++) acc += dA[w * E + c] * G[c];
                        d4[w] = acc;
                    }
                    double gs;
                    net_backward(en, ws, dG, &gs);
                    d4[0] += gs;
                }
            }
        }
        free(ws); free(A); free(dA); free(DR); free(dDR);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* D6: force and virial, CPU branch dp_dp.py:436-459                         */
/* ------------------------------------------------------------------------ */
/*
 * list [N][S] 1-based (0 empty); dE [N][S][4]; Ri_d [N][S][4][3];
 * image_dr [N][S][4]; out force [N+nghost][3], virial [9] (upper triangle
 * computed as dR_a * (dE.Ri_d)_c, lower copied from upper).
 */
void oracle_dp_force_virial(const int32_t *list, const double *dE, const double *Ri_d,
                            const double *image_dr, int natoms, int nghost, int S,
                            double *force, double *virial)
{
    memset(force, 0, sizeof(double) * 3 * (natoms + nghost));
    double v[9] = {0};
    for (int i = 0; i < natoms; i++) {
        for (int k = 0; k < S; k++) {
            const size_t s = (size_t)i * S + k;
            double f[3] = {0, 0, 0};
            for (int w = 0; w < 4; w++)
                for (int a = 0; a < 3; a++) f[a] += dE[4 * s + w] * Ri_d[12 * s + 3 * w + a];
            for (int a = 0; a < 3; a++) force[3 * i + a] -= f[a];
            const int j = list[s];
            if (j > 0)
                for (int a = 0; a < 3; a++) force[3 * (j - 1) + a] += f[a];
            const double *d = image_dr + 4 * s;
            v[0] += d[1] * f[0]; v[1] += d[1] * f[1]; v[2] += d[1] * f[2];
            v[4] += d[2] * f[1]; v[5] += d[2] * f[2]; v[8] += d[3] * f[2];
        }
    }
    v[3] = v[1]; v[6] = v[2]; v[7] = v[5];
    memcpy(virial, v, sizeof(v));
}

/* ------------------------------------------------------------------------ */
/* D7/D7b/D8: the CalcOps operator contract, on one flat batch               */
/* ------------------------------------------------------------------------ */
/*
 * calculateForce: nlist [B][N][S] 0-based (-1 empty); dE [B][N][S*4];
 * Ri_d [B][N][S*4][3]; F_center [B][N][3]; out force [B][N+nghost][3] =
 * center term copied into rows < N, plus dE.Ri_d scattered to row j.
 * (Rows are strided by N+nghost; the reference strides the scatter by N,
 * which differs only when nghost > 0 and B > 1.)
 */
void oracle_dp_force_op(const int32_t *nlist, const double *dE, const double *Ri_d,
                        const double *F_center, int B, int N, int S, int nghost, double *force)
{
    const int NA = N + nghost;
    memset(force, 0, sizeof(double) * B * NA * 3);
    for (int b = 0; b < B; b++) {
        for (int i = 0; i < N; i++)
            for (int a = 0; a < 3; a++) force[((size_t)b * NA + i) * 3 + a] = F_center[((size_t)b * N + i) * 3 + a];
        for (int i = 0; i < N; i++)
            for (int k = 0; k < S; k++) {
                const size_t s = ((size_t)b * N + i) * S + k;
                const int j = nlist[s];
                if (j < 0) continue;
                for (int a = 0; a < 3; a++) {
                    double acc = 0.0;
                    for (int w = 0; w < 4; w++) acc += dE[4 * s + w] * Ri_d[12 * s + 3 * w + a];
                    force[((size_t)b * NA + j) * 3 + a] += acc;
                }
            }
    }
}

/*
 * calculateVirial: atom_virial[b][j][3a+c] += sum_w dE_w Ri_d[w][a] Rij[c]
 * (calculateVirial.cu:78-85); virial[b][v] = sum over N+nghost rows.
 * Rij [B][N][S][3].
 */
void oracle_dp_virial_op(const int32_t *nlist, const double *dE, const double *Rij,
                         const double *Ri_d, int B, int N, int S, int nghost,
                         double *virial, double *atom_virial)
{
    const int NA = N + nghost;
    memset(atom_virial, 0, sizeof(double) * B * NA * 9);
    for (int b = 0; b < B; b++) {
        for (int i = 0; i < N; i++)
            for (int k = 0; k < S; k++) {
                const size_t s = ((size_t)b * N + i) * S + k;
                const int j = nlist[s];
                if (j < 0) continue;
                for (int v = 0; v < 9; v++) {
                    const int a = v / 3, c = v % 3;
                    double acc = 0.0;
                    for (int w = 0; w < 4; w++) acc += dE[4 * s + w] * Ri_d[12 * s + 3 * w + a] * Rij[3 * s + c];
                    atom_virial[((size_t)b * NA + j) * 9 + v] += acc;
                }
            }
        for (int v = 0; v < 9; v++) {
            double acc = 0.0;
            for (int j = 0; j < NA; j++) acc += atom_virial[((size_t)b * NA + j) * 9 + v];
            virial[b * 9 + v] = acc;
        }
    }
}

/*
 * Backward of calculateForce wrt dE (calculateForceGrad.cu:9-59):
 * g[b][i][s][w] = -gF[b][i].Ri_d[s][w] + gF[b][j mod N].Ri_d[s][w] (j >= 0).
 * gF [B][N+nghost][3] (only rows < N are read, as the reference folds j).
 */
void oracle_dp_force_grad_op(const int32_t *nlist, const double *Ri_d, const double *gF,
                             int B, int N, int S, int nghost, double *g)
{
    const int NA = N + nghost;
    for (int b = 0; b < B; b++)
        for (int i = 0; i < N; i++)
            for (int k = 0; k < S; k++) {
                const size_t s = ((size_t)b * N + i) * S + k;
                int j = nlist[s];
                const double *gi = gF + ((size_t)b * NA + i) * 3;
                for (int w = 0; w < 4; w++) {
                    const double *rd = Ri_d + 12 * s + 3 * w;
                    double acc = -(gi[0] * rd[0] + gi[1] * rd[1] + gi[2] * rd[2]);
                    if (j >= 0) {
                        if (j >= N) j = j % N;
                        const double *gj = gF + ((size_t)b * NA + j) * 3;
                        acc += gj[0] * rd[0] + gj[1] * rd[1] + gj[2] * rd[2];
                    }
                    g[4 * s + w] = acc;
                }
            }
}

/*
 * Backward of calculateVirial wrt dE (calculateVirialGrad.cu:15-66):
 * g[b][i][s][w] = sum_{a,c} gV[b][3a+c] Rij[s][c] Ri_d[s][w][a], 0 for j < 0.
 */
void oracle_dp_virial_grad_op(const int32_t *nlist, const double *Rij, const double *Ri_d,
                              const double *gV, int B, int N, int S, double *g)
{
    for (int b = 0; b < B; b++)
        for (int i = 0; i < N; i++)
            for (int k = 0; k < S; k++) {
                const size_t s = ((size_t)b * N + i) * S + k;
                for (int w = 0; w < 4; w++) {
                    double acc = 0.0;
                    if (nlist[s] >= 0)
                        for (int a = 0; a < 3; a++)
                            for (int c = 0; c < 3; c++)
                                acc += gV[b * 9 + 3 * a + c] * Rij[3 * s + c] * Ri_d[12 * s + 3 * w + a];
                    g[4 * s + w] = acc;
                }
            }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of later oracle calls (the CPU baseline times one core and all cores). */
void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
