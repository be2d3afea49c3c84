/*
 * FP64 CPU oracle for the NEP neighbor search (TEST INFRASTRUCTURE ONLY; the
 * checker for matpl_amd's HIP kernels, never part of the product).
 *
 *   oracle_nep_expand_box  src/pre_data/nep_data_loader.py:220-254
 *   oracle_nep_neighbor    src/op/src/cpu_calculate_nepneighbor.cpp:97-203
 *                          (= src/op/kernel/calculateNepNeighbor.cu:166-281)
 *   oracle_nep_maxneigh    src/op/src/cpu_calculate_nepneighbor.cpp:6-95
 *
 * Frames are packed along the atom axis; every atom of a frame is tested
 * against every atom of the same frame in every replica (ia, ib, ic) in
 * [0, num_cell), self excluded in replica 0, with the minimum image taken in
 * the num_cell-expanded box, in the order (n2, ia, ib, ic). Neighbor indices
 * are global (packed) and the lists are padded with -1. Compiled with
 * -ffp-contract=off so the FP64 arithmetic rounds as written.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* lattice_cm: lattice transposed and flattened (column-major vectors), as
 * nep_data_loader passes lattice.T.flatten(). Returns the volume. */
double oracle_nep_expand_box(const double *lat, double cutoff, int64_t *num_cell, double *box)
{
    const double a[3] = {lat[0], lat[3], lat[6]};
    const double b[3] = {lat[1], lat[4], lat[7]};
    const double c[3] = {lat[2], lat[5], lat[8]};
    double det = lat[0] * (lat[4] * lat[8] - lat[5] * lat[7]) - lat[1] * (lat[3] * lat[8] - lat[5] * lat[6]) +
                 lat[2] * (lat[3] * lat[7] - lat[4] * lat[6]);
    const double volume = fabs(det);
#define AREA(u, v)                                                                                     \
    sqrt(((u)[1] * (v)[2] - (u)[2] * (v)[1]) * ((u)[1] * (v)[2] - (u)[2] * (v)[1]) +                  \
         ((u)[2] * (v)[0] - (u)[0] * (v)[2]) * ((u)[2] * (v)[0] - (u)[0] * (v)[2]) +                  \
         ((u)[0] * (v)[1] - (u)[1] * (v)[0]) * ((u)[0] * (v)[1] - (u)[1] * (v)[0]))
    num_cell[0] = (int64_t)ceil(2.0 * cutoff / (volume / AREA(b, c)));
    num_cell[1] = (int64_t)ceil(2.0 * cutoff / (volume / AREA(c, a)));
    num_cell[2] = (int64_t)ceil(2.0 * cutoff / (volume / AREA(a, b)));
#undef AREA
    for (int k = 0; k < 3; k++) {
        box[3 * k + 0] = lat[3 * k + 0] * (double)num_cell[0];
        box[3 * k + 1] = lat[3 * k + 1] * (double)num_cell[1];
        box[3 * k + 2] = lat[3 * k + 2] * (double)num_cell[2];
    }
    box[9] = box[4] * box[8] - box[5] * box[7];
    box[10] = box[2] * box[7] - box[1] * box[8];
    box[11] = box[1] * box[5] - box[2] * box[4];
    box[12] = box[5] * box[6] - box[3] * box[8];
    box[13] = box[0] * box[8] - box[2] * box[6];
    box[14] = box[2] * box[3] - box[0] * box[5];
    box[15] = box[3] * box[7] - box[4] * box[6];
    box[16] = box[1] * box[6] - box[0] * box[7];
    box[17] = box[0] * box[4] - box[1] * box[3];
    det *= (double)(num_cell[0] * num_cell[1] * num_cell[2]);
    for (int n = 9; n < 18; n++) box[n] /= det;
    return volume;
}

static void apply_mic(const double *h, double *x, double *y, double *z)
{
    double sx = h[9] * *x + h[10] * *y + h[11] * *z;
    double sy = h[12] * *x + h[13] * *y + h[14] * *z;
    double sz = h[15] * *x + h[16] * *y + h[17] * *z;
    sx -= nearbyint(sx);
    sy -= nearbyint(sy);
    sz -= nearbyint(sz);
    *x = h[0] * sx + h[1] * sy + h[2] * sz;
    *y = h[3] * sx + h[4] * sy + h[5] * sz;
    *z = h[6] * sx + h[7] * sy + h[8] * sz;
}

/* Visits every (n1, n2, replica) pair in reference order; mode 0 = neighbor
 * lists, 1 = counts, 2 = per-type counts. Returns 0, or 1 + an overflowing
 * atom (lists only). */
static int nep_scan(int mode, int64_t nframes, const int64_t *num_atoms, const int64_t *type_map, int ntypes,
                    const double *box, const double *box_orig, const int64_t *num_cell, const double *pos,
                    double rc2b, double rc3b, int64_t maxr, int64_t maxa, int64_t *NN_r, int64_t *NL_r,
                    int64_t *NN_a, int64_t *NL_a, double *Ri_r, double *Ri_a)
{
    int64_t N1 = 0;
    int err = 0;
    for (int64_t f = 0; f < nframes; f++) {
        const int64_t N2 = N1 + num_atoms[f];
        const double *h = box + 18 * f, *bo = box_orig + 9 * f;
        const int64_t *nc = num_cell + 3 * f;
        for (int64_t n1 = N1; n1 < N2; n1++) {
            const double x1 = pos[3 * n1], y1 = pos[3 * n1 + 1], z1 = pos[3 * n1 + 2];
            int64_t cr = 0, ca = 0;
            int64_t crt[64], cat[64];
            for (int t = 0; t < ntypes && t < 64; t++) crt[t] = cat[t] = 0;
            for (int64_t n2 = N1; n2 < N2; n2++) {
                for (int64_t ia = 0; ia < nc[0]; ia++)
                for (int64_t ib = 0; ib < nc[1]; ib++)
                for (int64_t ic = 0; ic < nc[2]; ic++) {
                    if (ia == 0 && ib == 0 && ic == 0 && n1 == n2) continue;
                    const double dx = bo[0] * (double)ia + bo[1] * (double)ib + bo[2] * (double)ic;
                    const double dy = bo[3] * (double)ia + bo[4] * (double)ib + bo[5] * (double)ic;
                    const double dz = bo[6] * (double)ia + bo[7] * (double)ib + bo[8] * (double)ic;
                    double x12 = pos[3 * n2] + dx - x1;
                    double y12 = pos[3 * n2 + 1] + dy - y1;
                    double z12 = pos[3 * n2 + 2] + dz - z1;
                    apply_mic(h, &x12, &y12, &z12);
                    const double d2 = x12 * x12 + y12 * y12 + z12 * z12;
                    if (mode == 2) {
                        const int t2 = (int)type_map[n2];
                        if (d2 < rc2b * rc2b) crt[t2]++;
                        if (d2 < rc3b * rc3b) cat[t2]++;
                        continue;
                    }
                    if (d2 < rc2b * rc2b) {
                        if (mode == 0) {
                            if (cr >= maxr) { if (!err) err = (int)n1 + 1; }
                            else {
                                NL_r[n1 * maxr + cr] = n2;
                                double *o = Ri_r + (n1 * maxr + cr) * 4;
                                o[0] = sqrt(d2); o[1] = x12; o[2] = y12; o[3] = z12;
                            }
                        }
                        cr++;
                    }
                    if (d2 < rc3b * rc3b) {
                        if (mode == 0) {
                            if (ca >= maxa) { if (!err) err = (int)n1 + 1; }
                            else {
                                NL_a[n1 * maxa + ca] = n2;
                                double *o = Ri_a + (n1 * maxa + ca) * 4;
                                o[0] = sqrt(d2); o[1] = x12; o[2] = y12; o[3] = z12;
                            }
                        }
                        ca++;
                    }
                }
            }
            if (mode == 2) {
                for (int t = 0; t < ntypes; t++) { NN_r[n1 * ntypes + t] = crt[t]; NN_a[n1 * ntypes + t] = cat[t]; }
            } else {
                NN_r[n1] = cr;
                NN_a[n1] = ca;
            }
        }
        N1 = N2;
    }
    return err;
}

/* calculate_neighbor(with_rij=True): NL padded with -1, Ri = [r, x, y, z] (0 padded). */
int oracle_nep_neighbor(int64_t nframes, const int64_t *num_atoms, const double *box, const double *box_orig,
                        const int64_t *num_cell, const double *pos, double rc2b, double rc3b, int64_t maxr,
                        int64_t maxa, int64_t ntot, int64_t *NN_r, int64_t *NL_r, int64_t *NN_a, int64_t *NL_a,
                        double *Ri_r, double *Ri_a)
{
    for (int64_t k = 0; k < ntot * maxr; k++) NL_r[k] = -1;
    for (int64_t k = 0; k < ntot * maxa; k++) NL_a[k] = -1;
    memset(Ri_r, 0, sizeof(double) * ntot * maxr * 4);
    memset(Ri_a, 0, sizeof(double) * ntot * maxa * 4);
    return nep_scan(0, nframes, num_atoms, NULL, 0, box, box_orig, num_cell, pos, rc2b, rc3b, maxr, maxa, NN_r,
                    NL_r, NN_a, NL_a, Ri_r, Ri_a);
}

/* calculate_maxneigh: per-atom counts ([N]) or per-atom per-type counts ([N][T]). */
void oracle_nep_maxneigh(int64_t nframes, const int64_t *num_atoms, const double *box, const double *box_orig,
                         const int64_t *num_cell, const double *pos, double rc2b, double rc3b, int ntypes,
                         const int64_t *type_map, int with_type, int64_t *NN_r, int64_t *NN_a)
{
    nep_scan(with_type ? 2 : 1, nframes, num_atoms, type_map, ntypes, box, box_orig, num_cell, pos, rc2b, rc3b, 0,
             0, NN_r, NULL, NN_a, NULL, NULL, NULL);
}
