// DP neighbor search on gfx950 (the NeighConst.find_neighbore contract,
// src/pre_data/gen_feature/NeighConst.f90:14-149).
//
// Output layout is the reference's, per frame-packed atom i:
//   list  [Ntot][T][m] int32, 1-based neighbor index within the frame, 0 = empty
//   dR    [Ntot][T][m][3] float64 = r_j - r_i (Cartesian)
// plus, for the fused path and MD,
//   counts[Ntot][T] int32 stored neighbors per bucket
//   shift [Ntot][T][m] int32, packed integer image (s1,s2,s3), each +128 in a byte,
//         such that dR = (f_j - f_i + s) . lattice with f the wrapped fractions.
//
// Membership and dR follow the reference bit for bit: positions are wrapped
// with fmod(x+10,1); a frame is "small" when any lattice vector is shorter than
// 2*rc_m, in which case every image in [-nperiod, nperiod]^3 of every atom is
// tested in (j, i1, i2, i3) order; otherwise each fractional difference takes
// the nearest of d, d+1, d-1. dR and |dR|^2 are evaluated in FP64 in the
// reference's operand order with FMA contraction disabled in this file.
// Buckets list neighbors in the reference's order (ascending j, then image).
//
// Two kernels:
//   brute  : one wavefront per atom, j in 64-wide chunks, integer wave scans to
//            place hits; any number of frames (batched training data).
//   cell   : one large frame; atoms binned into a fractional cell grid with a
//            counting sort, each wavefront scans the <=27 neighbor cells,
//            appends hits to per-type LDS buckets, then ranks them by j.
#include "common.h"

#pragma clang fp contract(off)

namespace matpl {
namespace {

constexpr int WAVE = 64;

struct Lattice {
    double a[9];   // row i = lattice vector i
};

__device__ __forceinline__ int pack_shift(int s1, int s2, int s3) {
    return (s1 + 128) | ((s2 + 128) << 8) | ((s3 + 128) << 16);
}

__device__ __forceinline__ void frac_to_cart(const double *L, double dx1, double dx2, double dx3,
                                             double &dx, double &dy, double &dz) {
    dx = L[0] * dx1 + L[3] * dx2 + L[6] * dx3;
    dy = L[1] * dx1 + L[4] * dx2 + L[7] * dx3;
    dz = L[2] * dx1 + L[5] * dx2 + L[8] * dx3;
}

// Nearest-of {d, d+1, d-1} as NeighConst.f90:98-108; returns the shift taken.
__device__ __forceinline__ int near_image(double &d) {
    int s = 0;
    if (fabs(d + 1.0) < fabs(d)) { d = d + 1.0; s += 1; }
    if (fabs(d - 1.0) < fabs(d)) { d = d - 1.0; s -= 1; }
    return s;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        int u = __shfl_up(v, off, WAVE);
        if (lane >= off) v += u;
    }
    return v;
}

// A non-finite coordinate has no cell and no image: status gets STATUS_NONFINITE (before any
// search kernel can record an overflow) and the coordinate is searched as 0, so that every index
// the search forms stays in range.
constexpr int32_t STATUS_NONFINITE = -4;

__global__ void wrap_kernel(const double *__restrict__ frac, double *__restrict__ frac_w, int64_t n3,
                            int32_t *__restrict__ status) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double x = frac[i];
        if (!isfinite(x)) {
            atomicCAS(status, 0, STATUS_NONFINITE);
            frac_w[i] = 0.0;
        } else {
            frac_w[i] = fmod(x + 10.0, 1.0);
        }
    }
}

// ----------------------------------------------------------------- brute force
// One wavefront per atom. Frame geometry is read from lattice[frame*9].
__global__ void __launch_bounds__(256)
brute_kernel(const double *__restrict__ lattice, const double *__restrict__ fw,
             const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame,
             const int32_t *__restrict__ type_map, const double *__restrict__ rc_type, int T, int m,
             double rc_m, int64_t ntot, int32_t *__restrict__ list, double *__restrict__ dR,
             int32_t *__restrict__ counts, int32_t *__restrict__ shift, int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t i = blockIdx.x * (int64_t)(blockDim.x / WAVE) + (threadIdx.x / WAVE);
    if (i >= ntot) return;
    const int f = atom_frame[i];
    const int64_t j0 = frame_start[f], j1 = frame_start[f + 1];
    const double *L = lattice + 9 * (int64_t)f;
    int np[3];
    bool small_box = false;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double len = sqrt(L[3 * d] * L[3 * d] + L[3 * d + 1] * L[3 * d + 1] + L[3 * d + 2] * L[3 * d + 2]);
        np[d] = (int)(rc_m / len) + 1;
        if (len < 2.0 * rc_m) small_box = true;
    }
    if (!small_box) np[0] = np[1] = np[2] = 0;
    const double rc2 = rc_m * rc_m;
    const double rci = rc_type[type_map[i]];
    const double rci2 = rci * rci;
    const double fi0 = fw[3 * i], fi1 = fw[3 * i + 1], fi2 = fw[3 * i + 2];
    const int64_t S = (int64_t)T * m;
    int32_t *li = list + i * S;
    double *di = dR ? dR + i * S * 3 : nullptr;
    int32_t *si = shift ? shift + i * S : nullptr;

    // running stored / counted (rc_m) totals per type; T <= 16
    int base[16], baseM[16];
    for (int t = 0; t < T; t++) base[t] = baseM[t] = 0;
    bool overflow = false;

    for (int64_t jc = j0; jc < j1; jc += WAVE) {
        const int64_t j = jc + lane;
        const bool valid = j < j1;
        const int tj = valid ? type_map[j] : 0;
        double fj0 = 0, fj1 = 0, fj2 = 0;
        if (valid) { fj0 = fw[3 * j]; fj1 = fw[3 * j + 1]; fj2 = fw[3 * j + 2]; }
        // pass 1: counts for this lane's j
        int c = 0, cM = 0;
        if (valid) {
            for (int i1 = -np[0]; i1 <= np[0]; i1++)
            for (int i2 = -np[1]; i2 <= np[1]; i2++)
            for (int i3 = -np[2]; i3 <= np[2]; i3++) {
                double dx1 = fj0 - fi0, dx2 = fj1 - fi1, dx3 = fj2 - fi2;
                if (small_box) { dx1 = dx1 + (double)i1; dx2 = dx2 + (double)i2; dx3 = dx3 + (double)i3; }
                else { near_image(dx1); near_image(dx2); near_image(dx3); }
                double dx, dy, dz;
                frac_to_cart(L, dx1, dx2, dx3, dx, dy, dz);
                const double dd = dx * dx + dy * dy + dz * dz;
                if (dd < rc2 && dd > 1e-8) {
                    cM++;
                    if (dd < rci2) c++;
                }
            }
        }
        // place this chunk's hits after the running totals, type by type
        int off = 0;
        for (int t = 0; t < T; t++) {
            const int v = (tj == t && valid) ? c : 0;
            const int vM = (tj == t && valid) ? cM : 0;
            const int incl = wave_incl_scan(v);
            const int inclM = wave_incl_scan(vM);
            const int tot = __shfl(incl, WAVE - 1, WAVE);
            const int totM = __shfl(inclM, WAVE - 1, WAVE);
            if (tj == t) off = base[t] + incl - v;
            base[t] += tot;
            baseM[t] += totM;
            if (baseM[t] > m) overflow = true;
        }
        if (overflow) break;
        // pass 2: write this lane's hits in image order
        if (valid && c > 0) {
            int k = off;
            for (int i1 = -np[0]; i1 <= np[0]; i1++)
            for (int i2 = -np[1]; i2 <= np[1]; i2++)
            for (int i3 = -np[2]; i3 <= np[2]; i3++) {
                double dx1 = fj0 - fi0, dx2 = fj1 - fi1, dx3 = fj2 - fi2;
                int s1 = i1, s2 = i2, s3 = i3;
                if (small_box) { dx1 = dx1 + (double)i1; dx2 = dx2 + (double)i2; dx3 = dx3 + (double)i3; }
                else { s1 = near_image(dx1); s2 = near_image(dx2); s3 = near_image(dx3); }
                double dx, dy, dz;
                frac_to_cart(L, dx1, dx2, dx3, dx, dy, dz);
                const double dd = dx * dx + dy * dy + dz * dz;
                if (dd < rc2 && dd > 1e-8 && dd < rci2) {
                    const int64_t slot = (int64_t)tj * m + k;
                    li[slot] = (int32_t)(j - j0) + 1;
                    if (di) { di[3 * slot] = dx; di[3 * slot + 1] = dy; di[3 * slot + 2] = dz; }
                    if (si) si[slot] = pack_shift(s1, s2, s3);
                    k++;
                }
            }
        }
    }
    if (overflow) {
        if (lane == 0) atomicCAS(status, 0, (int)(i + 1));
        for (int t = 0; t < T; t++) base[t] = 0;
    }
    // zero the unused tail of every bucket
    for (int t = 0; t < T; t++) {
        for (int k = base[t] + lane; k < m; k += WAVE) {
            const int64_t slot = (int64_t)t * m + k;
            li[slot] = 0;
            if (di) { di[3 * slot] = 0.0; di[3 * slot + 1] = 0.0; di[3 * slot + 2] = 0.0; }
            if (si) si[slot] = pack_shift(0, 0, 0);
        }
        if (counts && lane == 0) counts[i * T + t] = base[t];
    }
}

// ------------------------------------------------------------------ cell list
struct Grid {
    int nc[3];
    int noff[3];       // number of distinct offsets per dim (1..3)
    int off[3][3];     // distinct offsets per dim
};

__global__ void bin_kernel(const double *__restrict__ fw, int64_t n, Grid g, int32_t *__restrict__ atom_cell,
                           int32_t *__restrict__ cell_count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int v = (int)(fw[3 * i + d] * g.nc[d]);
            c[d] = v < 0 ? 0 : (v >= g.nc[d] ? g.nc[d] - 1 : v);
        }
        const int cell = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
        atom_cell[i] = cell;
        atomicAdd(cell_count + cell, 1);
    }
}

// Exclusive scan of cell_count into cell_start (ncells+1 entries), one block.
__global__ void __launch_bounds__(1024) scan_kernel(const int32_t *__restrict__ cnt, int32_t *__restrict__ start,
                                                    int32_t *__restrict__ cursor, int ncells) {
    __shared__ int32_t part[1024 / WAVE];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    for (int base = 0; base < ncells; base += 1024) {
        const int idx = base + threadIdx.x;
        const int v = idx < ncells ? cnt[idx] : 0;
        const int incl = wave_incl_scan(v);
        if (lane == WAVE - 1) part[wid] = incl;
        __syncthreads();
        int wofs = 0;
        for (int w = 0; w < wid; w++) wofs += part[w];
        const int excl = carry + wofs + incl - v;
        if (idx < ncells) { start[idx] = excl; cursor[idx] = excl; }
        __syncthreads();
        if (threadIdx.x == 1023) carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[ncells] = carry;
}

__global__ void fill_cells_kernel(const int32_t *__restrict__ atom_cell, int64_t n, int32_t *__restrict__ cursor,
                                  int32_t *__restrict__ cell_atoms) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        cell_atoms[atomicAdd(cursor + atom_cell[i], 1)] = (int32_t)i;
}

constexpr int CELL_WAVES = 4;   // wavefronts (atoms) per block

__global__ void __launch_bounds__(CELL_WAVES * WAVE)
cell_kernel(Lattice lat, Grid g, const double *__restrict__ fw, const int32_t *__restrict__ atom_cell,
            const int32_t *__restrict__ cell_start, const int32_t *__restrict__ cell_atoms,
            const int32_t *__restrict__ type_map, const double *__restrict__ rc_type, int T, int m, double rc_m,
            int64_t n, int32_t *__restrict__ list, double *__restrict__ dR, int32_t *__restrict__ counts,
            int32_t *__restrict__ shift, int32_t *__restrict__ status) {
    extern __shared__ int32_t smem[];
    // wid is the same on every lane: the atom, its cell and the cells' ranges stay scalar
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    int32_t *bucket = smem + wid * (T * m + 2 * T);   // [T][m] hit indices
    int32_t *cnt = bucket + T * m;                    // [T] stored
    int32_t *cntM = cnt + T;                          // [T] counted at rc_m
    const int64_t i = blockIdx.x * (int64_t)CELL_WAVES + wid;
    const bool active = i < n;
    for (int t = lane; t < 2 * T; t += WAVE) cnt[t] = 0;
    __syncthreads();
    const double *L = lat.a;
    const double rc2 = rc_m * rc_m;
    double fi0 = 0, fi1 = 0, fi2 = 0, rci2 = 0;
    int ci[3] = {0, 0, 0};
    if (active) {
        fi0 = fw[3 * i]; fi1 = fw[3 * i + 1]; fi2 = fw[3 * i + 2];
        const double rci = rc_type[type_map[i]];
        rci2 = rci * rci;
        int c = atom_cell[i];
        ci[2] = c % g.nc[2]; c /= g.nc[2];
        ci[1] = c % g.nc[1]; ci[0] = c / g.nc[1];
    }
    if (active) {
        // The z cells of one (x, y) column are contiguous in cell_atoms, so the column's 1-3 z
        // offsets become at most two runs of consecutive cells (the periodic wrap splits one), and
        // the lanes sweep each run in one pass rather than cell by cell (about 90 candidates per
        // run against about 30 per cell: fewer, fuller passes of the FP64 distance test).
        int zr0[3], zr1[3], nzr = 0;
        {
            int zs[3];
            const int nzo = g.noff[2];
            for (int c = 0; c < nzo; c++) zs[c] = (ci[2] + g.off[2][c] + g.nc[2]) % g.nc[2];
            for (int x = 1; x < nzo; x++)
                for (int y = x; y > 0 && zs[y - 1] > zs[y]; y--) {
                    const int tmp = zs[y]; zs[y] = zs[y - 1]; zs[y - 1] = tmp;
                }
            for (int c = 0; c < nzo; c++) {
                if (nzr > 0 && zs[c] == zr1[nzr - 1] + 1) zr1[nzr - 1] = zs[c];
                else { zr0[nzr] = zs[c]; zr1[nzr] = zs[c]; nzr++; }
            }
        }
        for (int a = 0; a < g.noff[0]; a++)
        for (int b = 0; b < g.noff[1]; b++)
        for (int c = 0; c < nzr; c++) {
            const int cx = (ci[0] + g.off[0][a] + g.nc[0]) % g.nc[0];
            const int cy = (ci[1] + g.off[1][b] + g.nc[1]) % g.nc[1];
            const int col = (cx * g.nc[1] + cy) * g.nc[2];
            const int s0 = cell_start[col + zr0[c]], s1 = cell_start[col + zr1[c] + 1];
            // two candidates per lane per pass, both passes' loads issued before either is used
            for (int p0 = s0; p0 < s1; p0 += 2 * WAVE) {
                int jj[2];
                double fj[2][3];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int p = p0 + u * WAVE + lane;
                    jj[u] = p < s1 ? cell_atoms[p] : -1;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int j = jj[u] < 0 ? 0 : jj[u];
                    fj[u][0] = fw[3 * j]; fj[u][1] = fw[3 * j + 1]; fj[u][2] = fw[3 * j + 2];
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int j = jj[u];
                    if (j < 0) continue;
                    double dx1 = fj[u][0] - fi0, dx2 = fj[u][1] - fi1, dx3 = fj[u][2] - fi2;
                    near_image(dx1); near_image(dx2); near_image(dx3);
                    double dx, dy, dz;
                    frac_to_cart(L, dx1, dx2, dx3, dx, dy, dz);
                    const double dd = dx * dx + dy * dy + dz * dz;
                    if (dd < rc2 && dd > 1e-8) {
                        const int tj = type_map[j];
                        atomicAdd(cntM + tj, 1);
                        if (dd < rci2) {
                            const int k = atomicAdd(cnt + tj, 1);
                            if (k < m) bucket[tj * m + k] = j;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (!active) return;
    bool overflow = false;
    for (int t = 0; t < T; t++) overflow |= cntM[t] > m;
    if (overflow) {
        if (lane == 0) atomicCAS(status, 0, (int)(i + 1));
        for (int t = 0; t < T; t++) cnt[t] = 0;
    }
    const int64_t S = (int64_t)T * m;
    int32_t *li = list + i * S;
    double *di = dR ? dR + i * S * 3 : nullptr;
    int32_t *si = shift ? shift + i * S : nullptr;
    for (int t = 0; t < T; t++) {
        const int nt = cnt[t];
        const int32_t *bk = bucket + t * m;
        for (int e = lane; e < nt; e += WAVE) {
            const int j = bk[e];
            int rank = 0;
            for (int q = 0; q < nt; q++) rank += bk[q] < j;
            double dx1 = fw[3 * j] - fi0, dx2 = fw[3 * j + 1] - fi1, dx3 = fw[3 * j + 2] - fi2;
            const int s1 = near_image(dx1), s2 = near_image(dx2), s3 = near_image(dx3);
            const int64_t slot = (int64_t)t * m + rank;
            li[slot] = j + 1;
            if (di) {
                double dx, dy, dz;
                frac_to_cart(L, dx1, dx2, dx3, dx, dy, dz);
                di[3 * slot] = dx; di[3 * slot + 1] = dy; di[3 * slot + 2] = dz;
            }
            if (si) si[slot] = pack_shift(s1, s2, s3);
        }
        for (int k = nt + lane; k < m; k += WAVE) {
            const int64_t slot = (int64_t)t * m + k;
            li[slot] = 0;
            if (di) { di[3 * slot] = 0.0; di[3 * slot + 1] = 0.0; di[3 * slot + 2] = 0.0; }
            if (si) si[slot] = pack_shift(0, 0, 0);
        }
        if (counts && lane == 0) counts[i * T + t] = nt;
    }
}

// Reverse slot of every used bucket entry: for s = (i, t, k) naming
// neighbor j with image shift sh, rev[s] is the entry of j's bucket of type
// t_i that names i with image -sh. The list is symmetric because dR_ji is the
// exact negation of dR_ij in the cutoff test; a missing partner (per-type
// cutoffs that differ, or a tie in the nearest-image rule) sets status.
//
// Each pair of entries is matched once, by its owner: the entry of the atom with the smaller
// index (for an atom's own images, the entry whose packed shift is the larger of the two). A
// first pass sets every slot to -1; the owner then binary-searches j's bucket (sorted by
// (j, image)) for its partner and writes both directions. Every load of the search touches a
// different cache line per lane, so the kernel is bound by the number of such loads, which the
// owner rule halves (a two-pass scan, every 16th entry then 15, took twice as long as the seven
// or eight probes of a binary search: 807 against 386 us on a 99,636-atom list of 160 slots per
// type, when every entry searched). An entry whose owner-side partner is missing sets status
// directly; one missing on the other side stays -1, which reverse_check_kernel (a pass over the
// used entries) turns into status.
constexpr int REV_WAVES = 4;

__global__ void __launch_bounds__(256)
reverse_init_kernel(int32_t *__restrict__ rev, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots; s += (int64_t)gridDim.x * blockDim.x)
        rev[s] = -1;
}

__global__ void __launch_bounds__(REV_WAVES * WAVE)
reverse_kernel(const int32_t *__restrict__ list, const int32_t *__restrict__ shift,
               const int32_t *__restrict__ counts, const int32_t *__restrict__ type_map,
               const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame, int T, int m,
               int64_t n, int32_t *__restrict__ rev, int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int64_t i = (int64_t)blockIdx.x * REV_WAVES + threadIdx.x / WAVE; i < n;
         i += (int64_t)gridDim.x * REV_WAVES) {
        const int64_t fb = frame_start ? frame_start[atom_frame[i]] : 0;
        const int ti = type_map[i];
        const int target = (int)(i - fb) + 1;
        int total = 0;
        for (int t = 0; t < T; t++) total += counts[i * T + t];
        for (int e0 = 0; e0 < total; e0 += WAVE) {
            const int e = e0 + lane;
            if (e < total) {
                int t = 0, k = e;
                while (k >= counts[i * T + t]) k -= counts[i * T + t++];
                const int64_t s = (i * T + t) * (int64_t)m + k;
                const int64_t j = fb + list[s] - 1;
                const int sh = shift[s];
                const int want = pack_shift(-((sh & 255) - 128), -(((sh >> 8) & 255) - 128),
                                            -(((sh >> 16) & 255) - 128));
                if (i < j || (i == j && sh > want)) {
                    const int64_t base = (j * T + ti) * (int64_t)m;
                    const int cnt = counts[j * T + ti];
                    int lo = 0, hi = cnt;          // first entry naming an atom >= i (buckets are sorted)
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (list[base + mid] < target) lo = mid + 1; else hi = mid;
                    }
                    int found = -1;
                    for (int q = lo; q < cnt && list[base + q] == target; q++)
                        if (shift[base + q] == want) { found = q; break; }
                    if (found < 0) {
                        atomicCAS(status, 0, (int)(i + 1));
                    } else {
                        rev[s] = (int32_t)(base + found);
                        rev[base + found] = (int32_t)s;
                    }
                }
            }
        }
    }
}

// status = 1 + i for the first atom i with a used entry left without its partner (the owner's list
// lacks it); one wavefront per atom over its used entries
__global__ void __launch_bounds__(REV_WAVES * WAVE)
reverse_check_kernel(const int32_t *__restrict__ counts, const int32_t *__restrict__ rev, int T, int m, int64_t n,
                     int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int64_t i = (int64_t)blockIdx.x * REV_WAVES + threadIdx.x / WAVE; i < n;
         i += (int64_t)gridDim.x * REV_WAVES) {
        bool bad = false;
        for (int t = 0; t < T; t++) {
            const int c = counts[i * T + t];
            for (int k = lane; k < c; k += WAVE) bad |= rev[(i * T + t) * (int64_t)m + k] < 0;
        }
        if (__ballot(bad) != 0ull && lane == 0) atomicCAS(status, 0, (int)(i + 1));
    }
}

// ------------------------------------------------------------------ LAMMPS packing
// PairMATPL::generate_neighdata (src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:522-616)
// on the device. Input is LAMMPS's full neighbor list as CSR (neighbors of atom i
// are neigh[firstneigh[i] .. firstneigh[i] + numneigh[i]), indices over local +
// ghost atoms) and positions x [nall][3] of local and ghost atoms. For list entry
// ii (centre i = ilist[ii]) every neighbor j with |x_j - x_i|^2 <= rc^2 goes,
// in LAMMPS list order, to bucket t_j = type_to_model[type[j] - 1] of row i:
//   list [inum][T][m] = j + 1 (0 = empty), dR [inum][T][m][4] = [r, dx, dy, dz],
//   imagetype[ii] = t_i, counts [inum][T] = neighbors found (may exceed m).
// Unused slots are zeroed (the reference keeps whatever the buffer held).
// status (zero before the call) receives the first problem: 1 + i for a bucket
// that overflowed m (the reference writes past it), -1 for a centre i outside
// [0, inum), -2 for a neighbor outside [0, nall) or a list range outside
// neigh[0 .. nneigh), -3 for a type with no model
// type. min_rsq (uint64, +inf bits before the call) receives the smallest
// accepted |dR|^2 as double bits (non-negative doubles order like their bits),
// for the reference's "two atoms too close" check (pair_matpl.cpp:605-611).
// One wavefront per list entry, lanes over its neighbors, integer wave scans
// to keep list order inside each bucket.
__global__ void __launch_bounds__(256)
lmp_pack_kernel(const double *__restrict__ x, const int32_t *__restrict__ type,
                const int32_t *__restrict__ type_to_model, int ntypes_lmp, const int32_t *__restrict__ ilist,
                const int32_t *__restrict__ numneigh, const int64_t *__restrict__ firstneigh,
                const int32_t *__restrict__ neigh, int64_t nneigh, int64_t inum, int64_t nall, int T, int m,
                double rc2,
                int32_t *__restrict__ imagetype, int32_t *__restrict__ list, double *__restrict__ dR,
                int32_t *__restrict__ counts, unsigned long long *__restrict__ min_rsq,
                int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t ii = blockIdx.x * (int64_t)(blockDim.x / WAVE) + (threadIdx.x / WAVE);
    if (ii >= inum) return;
    const int64_t i = ilist[ii];
    if (i < 0 || i >= inum) {
        if (lane == 0) atomicCAS(status, 0, -1);
        return;
    }
    auto model_type = [&](int64_t a) -> int {
        const int lt = type[a];
        return (lt >= 1 && lt <= ntypes_lmp) ? type_to_model[lt - 1] : -1;
    };
    const int ti = model_type(i);
    if (ti < 0 || ti >= T) {
        if (lane == 0) atomicCAS(status, 0, -3);
        return;
    }
    if (lane == 0) imagetype[ii] = ti;
    const int64_t S = (int64_t)T * m;
    int32_t *li = list + i * S;
    double *di = dR + i * S * 4;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    const int jnum = numneigh[i];
    const int64_t j0 = firstneigh[i];
    if (jnum < 0 || j0 < 0 || j0 + jnum > nneigh) {
        if (lane == 0) atomicCAS(status, 0, -2);
        return;
    }
    const int32_t *jl = neigh + j0;
    int base[16];
    for (int t = 0; t < T; t++) base[t] = 0;
    double rmin = __longlong_as_double(0x7ff0000000000000LL);
    bool bad_j = false, bad_t = false;
    for (int jc = 0; jc < jnum; jc += WAVE) {
        const int jj = jc + lane;
        int64_t j = jj < jnum ? (int64_t)jl[jj] : -1;
        if (jj < jnum && (j < 0 || j >= nall)) { bad_j = true; j = -1; }
        int tj = -1;
        double dx = 0.0, dy = 0.0, dz = 0.0, rsq = 0.0;
        if (j >= 0) {
            dx = x[3 * j] - xi;
            dy = x[3 * j + 1] - yi;
            dz = x[3 * j + 2] - zi;
            rsq = dx * dx + dy * dy + dz * dz;
            if (rsq <= rc2) {
                tj = model_type(j);
                if (tj < 0 || tj >= T) { bad_t = true; tj = -1; }
            }
        }
        int pos = 0;
        for (int t = 0; t < T; t++) {
            const int v = tj == t ? 1 : 0;
            const int incl = wave_incl_scan(v);
            if (tj == t) pos = base[t] + incl - 1;
            base[t] += __shfl(incl, WAVE - 1, WAVE);
        }
        if (tj >= 0) {
            rmin = fmin(rmin, rsq);
            if (pos < m) {
                const int64_t slot = (int64_t)tj * m + pos;
                li[slot] = (int32_t)(j + 1);
                di[4 * slot] = sqrt(rsq);
                di[4 * slot + 1] = dx;
                di[4 * slot + 2] = dy;
                di[4 * slot + 3] = dz;
            }
        }
    }
    bool over = false;
    for (int t = 0; t < T; t++) {
        over = over || base[t] > m;
        const int used = base[t] < m ? base[t] : m;
        for (int k = used + lane; k < m; k += WAVE) {
            const int64_t slot = (int64_t)t * m + k;
            li[slot] = 0;
            di[4 * slot] = di[4 * slot + 1] = di[4 * slot + 2] = di[4 * slot + 3] = 0.0;
        }
        if (counts && lane == 0) counts[i * T + t] = base[t];
    }
    // wave minimum of the accepted |dR|^2, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) rmin = fmin(rmin, __shfl_xor(rmin, off, WAVE));
    if (lane == 0 && min_rsq) atomicMin(min_rsq, (unsigned long long)__double_as_longlong(rmin));
    const bool any_bad_j = __any(bad_j), any_bad_t = __any(bad_t);
    if (lane == 0) {
        if (any_bad_j) atomicCAS(status, 0, -2);
        else if (any_bad_t) atomicCAS(status, 0, -3);
        else if (over) atomicCAS(status, 0, (int)(i + 1));
    }
}

bool small_box_host(const double *L, double rc) {
    for (int d = 0; d < 3; d++) {
        const double len = sqrt(L[3 * d] * L[3 * d] + L[3 * d + 1] * L[3 * d + 1] + L[3 * d + 2] * L[3 * d + 2]);
        if (len < 2.0 * rc) return true;
    }
    return false;
}

Grid make_grid(const double *L, double rc) {
    // plane spacing h_d = V / |a_e x a_f|
    const double *a = L, *b = L + 3, *c = L + 6;
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double bc[3], ca[3], ab[3];
    cross(b, c, bc); cross(c, a, ca); cross(a, b, ab);
    const double V = fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
    const double *ar[3] = {bc, ca, ab};
    Grid g;
    for (int d = 0; d < 3; d++) {
        const double area = sqrt(ar[d][0] * ar[d][0] + ar[d][1] * ar[d][1] + ar[d][2] * ar[d][2]);
        int nc = (int)floor(V / area / rc);
        if (nc < 1) nc = 1;
        if (nc > 1024) nc = 1024;
        g.nc[d] = nc;
        if (nc >= 3) { g.noff[d] = 3; g.off[d][0] = -1; g.off[d][1] = 0; g.off[d][2] = 1; }
        else if (nc == 2) { g.noff[d] = 2; g.off[d][0] = 0; g.off[d][1] = 1; g.off[d][2] = 0; }
        else { g.noff[d] = 1; g.off[d][0] = 0; g.off[d][1] = 0; g.off[d][2] = 0; }
    }
    return g;
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int64_t matpl_dp_neighbor_cell_workspace_bytes(const double *lattice_host, double rc_m, int64_t n) {
    Grid g = make_grid(lattice_host, rc_m);
    const int64_t ncells = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
    // atom_cell[n], cell_atoms[n], cell_count[ncells], cell_start[ncells+1], cursor[ncells]
    return (int64_t)sizeof(int32_t) * (2 * n + 3 * ncells + 1 + 64);
}

MATPL_EXPORT int matpl_dp_neighbor_is_small_box(const double *lattice_host, double rc_m) {
    return small_box_host(lattice_host, rc_m) ? 1 : 0;
}

MATPL_EXPORT int matpl_dp_neighbor_brute(const double *lattice, const double *frac, const int64_t *frame_start,
                                         const int32_t *atom_frame, const int32_t *type_map, const double *rc_type,
                                         int ntypes, int m, double rc_m, int64_t nframes, int64_t ntot,
                                         double *frac_wrapped, int32_t *list, double *dR, int32_t *counts,
                                         int32_t *shift, int32_t *status, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= 16, "dp_neighbor: ntypes must be in [1,16]");
    MATPL_CHECK_ARG(m >= 1 && nframes >= 0 && ntot >= 0 && rc_m > 0, "dp_neighbor: bad sizes");
    if (ntot == 0) return OK;
    MATPL_CHECK_ARG(lattice && frac && frame_start && atom_frame && type_map && rc_type && frac_wrapped &&
                        list && status, "dp_neighbor: null pointer");
    hipLaunchKernelGGL(wrap_kernel, dim3(grid_cap(ntot * 3)), dim3(256), 0, st, frac, frac_wrapped, ntot * 3,
                       status);
    int rc = check_launch("wrap_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(brute_kernel, dim3(ceil_div(ntot, 4)), dim3(256), 0, st, lattice, frac_wrapped, frame_start,
                       atom_frame, type_map, rc_type, ntypes, m, rc_m, ntot, list, dR, counts, shift, status);
    return check_launch("brute_kernel");
}

MATPL_EXPORT int matpl_dp_neighbor_cell(const double *lattice_host, const double *frac, const int32_t *type_map,
                                        const double *rc_type, int ntypes, int m, double rc_m, int64_t n,
                                        double *frac_wrapped, int32_t *list, double *dR, int32_t *counts,
                                        int32_t *shift, int32_t *status, void *workspace, int64_t workspace_bytes,
                                        void *stream) {
    hipStream_t st = (hipStream_t)stream;
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= 16, "dp_neighbor_cell: ntypes must be in [1,16]");
    MATPL_CHECK_ARG(m >= 1 && n >= 0 && rc_m > 0, "dp_neighbor_cell: bad sizes");
    MATPL_CHECK_ARG(lattice_host, "dp_neighbor_cell: null lattice");
    MATPL_CHECK_ARG(!small_box_host(lattice_host, rc_m),
                    "dp_neighbor_cell: a lattice vector is shorter than 2*rc; use the brute-force search");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(frac && type_map && rc_type && frac_wrapped && list && status && workspace,
                    "dp_neighbor_cell: null pointer");
    MATPL_CHECK_ARG(n < (int64_t)0x7fffffff, "dp_neighbor_cell: too many atoms");
    const int64_t need = matpl_dp_neighbor_cell_workspace_bytes(lattice_host, rc_m, n);
    MATPL_CHECK_ARG(workspace_bytes >= need, "dp_neighbor_cell: workspace %lld < %lld bytes",
                    (long long)workspace_bytes, (long long)need);
    const size_t lds = sizeof(int32_t) * CELL_WAVES * ((size_t)ntypes * m + 2 * ntypes);
    MATPL_CHECK_ARG(lds <= 64 * 1024, "dp_neighbor_cell: ntypes*m too large for LDS buckets");
    Grid g = make_grid(lattice_host, rc_m);
    const int ncells = g.nc[0] * g.nc[1] * g.nc[2];
    int32_t *ws = (int32_t *)workspace;
    int32_t *atom_cell = ws;
    int32_t *cell_atoms = atom_cell + n;
    int32_t *cell_count = cell_atoms + n;
    int32_t *cell_start = cell_count + ncells;
    int32_t *cursor = cell_start + ncells + 1;
    Lattice lat;
    for (int k = 0; k < 9; k++) lat.a[k] = lattice_host[k];
    int rc;
    hipLaunchKernelGGL(wrap_kernel, dim3(grid_cap(n * 3)), dim3(256), 0, st, frac, frac_wrapped, n * 3, status);
    if ((rc = check_launch("wrap_kernel"))) return rc;
    if ((rc = fill_async<int32_t>(cell_count, 0, ncells, st))) return rc;
    hipLaunchKernelGGL(bin_kernel, dim3(grid_cap(n)), dim3(256), 0, st, frac_wrapped, n, g, atom_cell, cell_count);
    if ((rc = check_launch("bin_kernel"))) return rc;
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, cell_count, cell_start, cursor, ncells);
    if ((rc = check_launch("scan_kernel"))) return rc;
    hipLaunchKernelGGL(fill_cells_kernel, dim3(grid_cap(n)), dim3(256), 0, st, atom_cell, n, cursor, cell_atoms);
    if ((rc = check_launch("fill_cells_kernel"))) return rc;
    hipLaunchKernelGGL(cell_kernel, dim3(ceil_div(n, CELL_WAVES)), dim3(CELL_WAVES * WAVE), lds, st, lat, g,
                       frac_wrapped, atom_cell, cell_start, cell_atoms, type_map, rc_type, ntypes, m, rc_m, n, list,
                       dR, counts, shift, status);
    return check_launch("cell_kernel");
}

MATPL_EXPORT int matpl_dp_neighbor_reverse(const int32_t *list, const int32_t *shift, const int32_t *counts,
                                           const int32_t *type_map, const int64_t *frame_start,
                                           const int32_t *atom_frame, int ntypes, int m, int64_t n, int32_t *rev,
                                           int32_t *status, void *stream) {
    MATPL_CHECK_ARG(ntypes >= 1 && m >= 1 && n >= 0, "dp_neighbor_reverse: bad sizes");
    MATPL_CHECK_ARG(n * ntypes * (int64_t)m < (int64_t)0x7fffffff, "dp_neighbor_reverse: too many slots for int32");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(list && shift && counts && type_map && rev && status, "dp_neighbor_reverse: null pointer");
    MATPL_CHECK_ARG((frame_start == nullptr) == (atom_frame == nullptr),
                    "dp_neighbor_reverse: frame_start and atom_frame go together");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nslots = n * ntypes * (int64_t)m;
    hipLaunchKernelGGL(reverse_init_kernel, dim3(grid_cap(nslots)), dim3(256), 0, st, rev, nslots);
    int rc = check_launch("reverse_init_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(reverse_kernel, dim3(grid_cap(n, REV_WAVES)), dim3(REV_WAVES * WAVE), 0, st, list, shift, counts,
                       type_map, frame_start, atom_frame, ntypes, m, n, rev, status);
    if ((rc = check_launch("reverse_kernel"))) return rc;
    hipLaunchKernelGGL(reverse_check_kernel, dim3(grid_cap(n, REV_WAVES)), dim3(REV_WAVES * WAVE), 0, st, counts, rev,
                       ntypes, m, n, status);
    return check_launch("reverse_check_kernel");
}

MATPL_EXPORT int matpl_dp_lmp_neighdata(const double *x, const int32_t *type, const int32_t *type_to_model,
                                        int ntypes_lmp, const int32_t *ilist, const int32_t *numneigh,
                                        const int64_t *firstneigh, const int32_t *neigh, int64_t nneigh, int64_t inum,
                                        int64_t nall,
                                        int ntypes, int m, double rcut, int32_t *imagetype, int32_t *list,
                                        double *dR, int32_t *counts, void *min_rsq, int32_t *status,
                                        void *stream) {
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= 16 && m >= 1 && inum >= 0 && nall >= inum && ntypes_lmp >= 1,
                    "dp_lmp_neighdata: bad sizes");
    MATPL_CHECK_ARG(inum * ntypes * (int64_t)m < (int64_t)0x7fffffff, "dp_lmp_neighdata: too many slots for int32");
    MATPL_CHECK_ARG(nall < (int64_t)0x7fffffff, "dp_lmp_neighdata: too many atoms for int32 indices");
    MATPL_CHECK_ARG(rcut > 0.0, "dp_lmp_neighdata: rcut must be positive");
    if (inum == 0) return OK;
    MATPL_CHECK_ARG(nneigh >= 0, "dp_lmp_neighdata: bad neighbor array size");
    MATPL_CHECK_ARG(x && type && type_to_model && ilist && numneigh && firstneigh && (neigh || nneigh == 0) &&
                        imagetype && list && dR && status,
                    "dp_lmp_neighdata: null pointer");
    hipLaunchKernelGGL(lmp_pack_kernel, dim3((unsigned)((inum + 3) / 4)), dim3(256), 0, (hipStream_t)stream, x, type,
                       type_to_model, ntypes_lmp, ilist, numneigh, firstneigh, neigh, nneigh, inum, nall, ntypes, m,
                       rcut * rcut, imagetype, list, dR, counts, (unsigned long long *)min_rsq, status);
    return check_launch("lmp_pack_kernel");
}
