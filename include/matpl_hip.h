/*
 * matpl_hip.h — C ABI of the MI355X (gfx950) descriptor-and-force library
 * (libmatpl_hip.so).
 *
 * Every entry point takes plain device pointers and sizes, enqueues its work
 * on `stream` (a hipStream_t passed as void*, NULL = default stream) and
 * returns a status code: 0 on success, MATPL_ERR_* otherwise, with a message
 * in matpl_last_error(). Nothing here synchronizes the stream or allocates
 * device memory; entry points that need scratch take a caller workspace
 * sized by the matching *_workspace_* query.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the MatPL source tree).
 */
#ifndef MATPL_HIP_H
#define MATPL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MATPL_OK 0
#define MATPL_ERR_ARGUMENT -1  /* bad size, null or misaligned pointer */
#define MATPL_ERR_HIP -2       /* a HIP runtime call or launch failed */
#define MATPL_ERR_OVERFLOW -3  /* a neighbor bucket overflowed */

const char *matpl_last_error(void);
int matpl_abi_version(void);

/* ------------------------------------------------------------------ DP ops */
/*
 * calculateForce: force[b][0:N] = F_center[b]; force[b][j] += dE.Ri_d for
 * every slot with nlist >= 0.  nlist [B][N][S] int32 0-based (-1 empty),
 * dE [B][N][S*4], Ri_d [B][N][S*4][3], F_center [B][N][3],
 * force out [B][N+nghost][3]. dE and Ri_d 16-byte aligned.
 * Replaces CalculateForceFuncs::forward (src/op/src/CalcOps.cpp:6-23) ->
 * torch_launch_calculate_force (src/op/src/calculate_force.cpp:4-41) ->
 * force_calc (src/op/kernel/calculateForce.cu:49-142).
 */
int matpl_dp_force_f32(const int32_t *nlist, const float *dE, const float *Ri_d, const float *F_center,
                       int64_t B, int64_t N, int64_t S, int64_t nghost, float *force, void *stream);
int matpl_dp_force_f64(const int32_t *nlist, const double *dE, const double *Ri_d, const double *F_center,
                       int64_t B, int64_t N, int64_t S, int64_t nghost, double *force, void *stream);

/*
 * calculateVirial: atom_virial[b][j][3a+c] = sum over slots pointing at j of
 * (dE.Ri_d)_a * Rij_c; virial[b][v] = sum_j atom_virial[b][j][v].
 * Rij [B][N][S][3]; workspace of matpl_dp_virial_workspace_elems(B,N,nghost)
 * elements of the same type.
 * Replaces CalculateVirialFuncs::forward (src/op/src/CalcOps.cpp:170-186) ->
 * torch_launch_calculate_virial_force (src/op/src/calculate_virial_force.cpp:4-48)
 * -> virial_deriv_wrt_neighbors_a + atom_virial_reduction
 * (src/op/kernel/calculateVirial.cu:18-120).
 */
int matpl_dp_virial_workspace_elems(int64_t B, int64_t N, int64_t nghost);
int matpl_dp_virial_f32(const int32_t *nlist, const float *dE, const float *Rij, const float *Ri_d, int64_t B,
                        int64_t N, int64_t S, int64_t nghost, float *virial, float *atom_virial, float *workspace,
                        int64_t workspace_elems, void *stream);
int matpl_dp_virial_f64(const int32_t *nlist, const double *dE, const double *Rij, const double *Ri_d, int64_t B,
                        int64_t N, int64_t S, int64_t nghost, double *virial, double *atom_virial,
                        double *workspace, int64_t workspace_elems, void *stream);

/*
 * Backward of calculateForce wrt dE: g[b][i][s][w] = -gF[b][i].Ri_d[s][w]
 * + gF[b][j mod N].Ri_d[s][w] (j >= 0). gF [B][N+nghost][3], g [B][N][S*4].
 * Replaces CalculateForceFuncs::backward (src/op/src/CalcOps.cpp:25-39) ->
 * force_grad_wrt_center_atom / force_grad_wrt_neighbors_a
 * (src/op/kernel/calculateForceGrad.cu:9-90).
 */
int matpl_dp_force_grad_f32(const int32_t *nlist, const float *Ri_d, const float *gF, int64_t B, int64_t N,
                            int64_t S, int64_t nghost, float *g, void *stream);
int matpl_dp_force_grad_f64(const int32_t *nlist, const double *Ri_d, const double *gF, int64_t B, int64_t N,
                            int64_t S, int64_t nghost, double *g, void *stream);

/*
 * Backward of calculateVirial wrt dE: g[b][i][s][w] = sum_{a,c} gV[b][3a+c]
 * Rij[s][c] Ri_d[s][w][a] (0 on empty slots). gV [B][9].
 * Replaces CalculateVirialFuncs::backward (src/op/src/CalcOps.cpp:188-202) ->
 * virial_grad_wrt_neighbors_a (src/op/kernel/calculateVirialGrad.cu:15-92).
 */
int matpl_dp_virial_grad_f32(const int32_t *nlist, const float *Rij, const float *Ri_d, const float *gV,
                             int64_t B, int64_t N, int64_t S, float *g, void *stream);
int matpl_dp_virial_grad_f64(const int32_t *nlist, const double *Rij, const double *Ri_d, const double *gV,
                             int64_t B, int64_t N, int64_t S, double *g, void *stream);

/* ---------------------------------------------------------- DP neighbors */
/*
 * Neighbor lists in the find_neighbore layout: list [Ntot][T][m] int32
 * (1-based index within the frame, 0 empty), dR [Ntot][T][m][3] float64
 * (r_j - r_i; may be NULL), counts [Ntot][T] (may be NULL), shift
 * [Ntot][T][m] packed image (may be NULL), frac_wrapped [Ntot][3] out.
 * status (device int32, zero before the call) receives 1 + the first atom
 * whose bucket overflowed m, or -4 if a coordinate is NaN or infinite (the
 * search then runs with that coordinate at 0). Replaces neighconst.find_neighbore
 * (src/pre_data/gen_feature/NeighConst.f90:14-149) as called by
 * find_neighbore (src/pre_data/dpuni_data_loader.py:595-650).
 *
 * _brute: any number of frames (lattice [F][9] device, frame_start [F+1]
 * int64 device, atom_frame [Ntot] int32 device), any box size.
 * _cell : one frame whose lattice vectors are all >= 2*rc_m (lattice host
 * [9]); workspace of matpl_dp_neighbor_cell_workspace_bytes bytes.
 */
int matpl_dp_neighbor_is_small_box(const double *lattice_host, double rc_m);
int64_t matpl_dp_neighbor_cell_workspace_bytes(const double *lattice_host, double rc_m, int64_t n);
int matpl_dp_neighbor_brute(const double *lattice, const double *frac, const int64_t *frame_start,
                            const int32_t *atom_frame, const int32_t *type_map, const double *rc_type, int ntypes,
                            int m, double rc_m, int64_t nframes, int64_t ntot, double *frac_wrapped, int32_t *list,
                            double *dR, int32_t *counts, int32_t *shift, int32_t *status, void *stream);
int matpl_dp_neighbor_cell(const double *lattice_host, const double *frac, const int32_t *type_map,
                           const double *rc_type, int ntypes, int m, double rc_m, int64_t n, double *frac_wrapped,
                           int32_t *list, double *dR, int32_t *counts, int32_t *shift, int32_t *status,
                           void *workspace, int64_t workspace_bytes, void *stream);

/*
 * DP neighbor data from a LAMMPS full neighbor list, on the device. Replaces
 * PairMATPL::generate_neighdata (src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:522-616).
 * Inputs: x [nall][3] float64 (local then ghost atoms), type [nall] int32
 * (LAMMPS types, 1-based), type_to_model [ntypes_lmp] int32 (model type
 * index per LAMMPS type, -1 = not in the model), ilist [inum], numneigh
 * [inum] and firstneigh [inum] int64 (offsets into neigh [nneigh], int32
 * atom indices < nall), all indexed by atom as in LAMMPS. For entry ii, centre
 * i = ilist[ii]: every neighbor j with |x_j - x_i|^2 <= rcut^2, in list
 * order, goes to bucket type_to_model[type[j]-1] of row i:
 * list [inum][ntypes][m] int32 = j + 1 (0 = empty), dR [inum][ntypes][m][4]
 * float64 = [r, dx, dy, dz], imagetype [inum] = model type of i, counts
 * [inum][ntypes] (may be NULL) = neighbors found. min_rsq (may be NULL;
 * device uint64 holding +inf double bits before the call) receives the
 * smallest accepted |dR|^2 as double bits. status (device int32, zero
 * before the call): 1 + i for a bucket over m, -1 centre index outside
 * [0, inum), -2 neighbor index outside [0, nall) or list range outside
 * neigh, -3 unmapped type.
 */
int matpl_dp_lmp_neighdata(const double *x, const int32_t *type, const int32_t *type_to_model, int ntypes_lmp,
                           const int32_t *ilist, const int32_t *numneigh, const int64_t *firstneigh,
                           const int32_t *neigh, int64_t nneigh, int64_t inum, int64_t nall, int ntypes, int m,
                           double rcut,
                           int32_t *imagetype, int32_t *list, double *dR, int32_t *counts, void *min_rsq,
                           int32_t *status, void *stream);

/*
 * Reverse slot map of a symmetric neighbor list: rev [N][T][m] int32, for a
 * used entry s of atom i naming j with image sh, the index of the entry of
 * j (bucket of i's type) naming i with image -sh; -1 on unused entries.
 * status receives 1 + an atom whose partner entry is missing. No reference
 * counterpart (the reference scatters with atomics, calculateForce.cu:92).
 */
int matpl_dp_neighbor_reverse(const int32_t *list, const int32_t *shift, const int32_t *counts,
                              const int32_t *type_map, const int64_t *frame_start, const int32_t *atom_frame,
                              int ntypes, int m, int64_t n, int32_t *rev, int32_t *status, void *stream);

/* ------------------------------------------------- DP environment matrix */
/*
 * Ri [B][N][S][4] and Ri_d [B][N][S][4][3] from ImageDR [B][N][S][4]
 * = (r, dx, dy, dz); type_map [N]; davg/dstd [T][S*4].
 * Replaces DP.calculate_Ri (src/model/dp_dp.py:217-302), which the
 * reference computes with torch tensor code.
 */
int matpl_dp_env_mat_f32(const float *image_dr, const int32_t *type_map, const float *davg, const float *dstd,
                         int64_t B, int64_t N, int64_t S, int ntypes, double rmin, double rmax, float *Ri,
                         float *Ri_d, void *stream);
int matpl_dp_env_mat_f64(const double *image_dr, const int32_t *type_map, const double *davg, const double *dstd,
                         int64_t B, int64_t N, int64_t S, int ntypes, double rmin, double rmax, double *Ri,
                         double *Ri_d, void *stream);

/*
 * davg/dstd [T][4] from list [N][S] (1-based) and dR [N][S][3].
 * Replaces calc_stat + smooth + compute_std
 * (src/pre_data/dpuni_data_loader.py:422-560).
 */
int64_t matpl_dp_calc_stat_workspace_elems(int ntypes);
int matpl_dp_calc_stat(const int32_t *list, const double *dR, const int32_t *type_map, int64_t N, int64_t S,
                       int ntypes, double rmin, double rmax, double *davg, double *dstd, double *workspace,
                       int64_t workspace_elems, void *stream);

/* ------------------------------------------- batched small products */
/*
 * Batched products of small matrices (float32 or float64, contiguous
 * operands), one per batch element n:
 *   mode 0: out[n][a][b] = sum_k P1[n][k][a] P2[n][k][b]   (P1^T P2)
 *   mode 1: out[n][K][a] = sum_j P1[n][k][j] P2[n][a][j]   (P1 P2^T, P1 [K][b], P2 [a][b])
 *   mode 2: out[n][K][b] = sum_i P1[n][k][i] P2[n][i][b]   (P1 P2,   P1 [K][a], P2 [a][b])
 * The two operands of one element must fit in 64 KB. No counterpart in the
 * reference: it replaces the batched torch.matmul calls of the per-atom
 * descriptor contraction and of DR (src/model/dp_dp.py:332, 369, 412) and
 * of their derivatives in the batched training step.
 */
int matpl_bsmm(int f64, int mode, const void *P1, const void *P2, void *out, int64_t n, int K, int a, int b,
               void *stream);

/* ------------------------------------------- fused DP descriptor + nets */
/*
 * Forward: for every atom, the in-cutoff neighbors of its neighbor list
 * (list/shift [N][T][mnl], counts [N][T], positions pos [N][3] float64,
 * lattice [F][9] device, atom_frame [N] / frame_start [F+1] or NULL for one
 * frame) give the normalized Ri (FP32), the embedding net output G on the
 * matrix cores (bf16 != 0: BF16 in / FP32 accumulate; 0: FP32), and
 * A[i] = scale * (sum Ri^T G + empty-slot term) [N][4][E], E in {16,25,32}.
 * frame_types [F] (uint32, or NULL): bit t set when type t is in the frame's
 * atom_type list. Only those types' buckets add an empty-slot term, and the
 * frame's scale is 1 / (m * popcount) (dp_dp.py:150-165, 197, 413); every
 * type of a frame's atoms must be in its list, and no list may be empty.
 * With NULL every model type takes part and scale (= 1 / (m * ntypes)) is
 * used.
 * vec/mat/tinfo are the packed net parameters (matpl_amd/dp_fused.py).
 * Backward: given dA = dEtot/dA, force [N][3] (float) receives -dE/dx and
 * virial [F][9] (float64, accumulated: zero it first) the
 * calculateVirial-convention virial. With rev (from
 * matpl_dp_neighbor_reverse) the neighbor forces are gathered without
 * atomics: the backward writes the force of each slot (i -> j) into j's
 * entry naming i, fslot [N][T][mnl][4] float (16-byte aligned), and the
 * centre terms fcenter [N][3], and matpl_dp_fused_gather then overwrites
 * force; with rev == NULL the backward adds them to force with atomics
 * (force zeroed first).
 * Together they replace DP.calculate_Ri + calculate_xyz_scater + the
 * autograd dE/dRi + calculateForce/calculateVirial of the reference
 * (src/model/dp_dp.py:217-469, src/model/dp_embedding.py:40-94); the fitting
 * net between them is evaluated by the caller.
 * status (device int32, zero before the call) receives 1 + the first atom
 * with more than m in-cutoff neighbors of one type.
 */
int matpl_dp_fused_forward(const double *pos, const double *lattice, const int32_t *atom_frame,
                           const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift,
                           const int32_t *counts, const int32_t *type_map, const float *vec, const void *mat,
                           const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E, int bf16,
                           double rcut, double rcut_smooth, double scale, float *A, int32_t *status, void *stream);
int matpl_dp_fused_backward(const double *pos, const double *lattice, const int32_t *atom_frame,
                            const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift,
                            const int32_t *counts, const int32_t *type_map, const float *vec, const void *mat,
                            const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E, int bf16,
                            double rcut, double rcut_smooth, double scale, const float *dA, const int32_t *rev,
                            float *fslot, float *fcenter, float *force, double *virial, int32_t *status,
                            void *stream);
/*
 * Store mode of the fused pair: matpl_dp_fused_forward_store computes A as
 * matpl_dp_fused_forward and also writes, for every in-cutoff neighbor slot,
 * the embedding output G(s) and its derivative dG/ds into gbuf
 * (matpl_dp_fused_gstore_words(natoms, ntypes, mnl, m, E, bf16) uint32 words,
 * 16-byte aligned; bf16 != 0: fp16 pairs, else float); the backward
 * matpl_dp_fused_backward_stored reads them instead of evaluating the net
 * again. Same results as the recomputing pair up to the fp16 rounding of G
 * and dG/ds in BF16 mode (FP32 mode: identical arithmetic). No counterpart in
 * the reference, whose autograd keeps the embedding activations for its
 * backward (dp_dp.py:429) the way this buffer does.
 */
int64_t matpl_dp_fused_gstore_words(int64_t natoms, int ntypes, int mnl, int m, int E, int bf16);
int matpl_dp_fused_forward_store(const double *pos, const double *lattice, const int32_t *atom_frame,
                                 const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                                 const int32_t *shift, const int32_t *counts, const int32_t *type_map,
                                 const float *vec, const void *mat, const float *tinfo, int64_t natoms, int ntypes,
                                 int mnl, int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                 float *A, uint32_t *gbuf, int32_t *status, void *stream);
int matpl_dp_fused_backward_stored(const double *pos, const double *lattice, const int32_t *atom_frame,
                                   const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                                   const int32_t *shift, const int32_t *counts, const int32_t *type_map,
                                   const float *vec, const void *mat, const float *tinfo, int64_t natoms, int ntypes,
                                   int mnl, int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                   const float *dA, const uint32_t *gbuf, const int32_t *rev, float *fslot,
                                   float *fcenter, float *force, double *virial, int32_t *status, void *stream);
/*
 * The fused forward / backward with a compressed embedding: G(s) and dG/ds
 * from piecewise polynomials instead of the net (the reference's DP compress,
 * dp_dp.py:350-355, 572-601 with the table of src/user/ckpt_compress.py).
 * ctab [T*T][rows][32][order+1] float (16-byte aligned): for pair p, row k
 * and feature f, the coefficients (highest power first) of the interval's
 * polynomial in x = s - knot_k; features E..31 hold zeros. Knots: x0 + k dx for k <= n_fine, then steps of 10 dx; s below
 * x0 or past the last row extrapolates the first or last interval.
 * order 3 or 5. mat is not read; vec supplies the empty-slot vector
 * G(s_pad), which the caller fills from the same table. Other arguments as
 * matpl_dp_fused_forward / matpl_dp_fused_backward.
 */
int matpl_dp_fused_forward_compressed(const double *pos, const double *lattice, const int32_t *atom_frame,
                                      const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift,
                                      const int32_t *counts, const int32_t *type_map, const float *vec,
                                      const float *ctab, int order, int rows, int n_fine, double x0, double dx,
                                      const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E, int bf16,
                                      double rcut, double rcut_smooth, double scale, float *A, int32_t *status,
                                      void *stream);
int matpl_dp_fused_backward_compressed(const double *pos, const double *lattice, const int32_t *atom_frame,
                                       const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift,
                                       const int32_t *counts, const int32_t *type_map, const float *vec,
                                       const float *ctab, int order, int rows, int n_fine, double x0, double dx,
                                       const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E,
                                       int bf16, double rcut, double rcut_smooth, double scale, const float *dA,
                                       const int32_t *rev, float *fslot, float *fcenter, float *force,
                                       double *virial, int32_t *status, void *stream);
/*
 * Force-loss double backward of the embedding nets for training (the
 * reference's dE = autograd.grad(Etot, Ri, create_graph=True) then
 * loss.backward(), src/model/dp_dp.py:429 and
 * src/PWMLFF/dp_mods/dp_trainer.py:274, through the embedding nets).
 * Geometry, weights and frame arguments as matpl_dp_fused_forward; the
 * embedding must have no resnet_dt and E in {16, 25}.
 * matpl_dp_train_tangent: dA [N][4][E] = c sum_j (v_j (x) G_j + Ri_j (x) v_j0
 * G'_j), the tangent of A along the force residual field resid [N][3]
 * (v_j = dRi_j/d(dR_j) (resid_j - resid_i), G' = dG/ds), and the in-cutoff
 * neighbor counts nin [N][T] (int32, may be NULL).
 * matpl_dp_train_embgrad: adds to grad [T*T][matpl_dp_train_grad_size(E)]
 * (float, zero it first; per pair W3 [E][E], b3, W2 [E][E], b2, w1, b1) the
 * gradient of sum_ij a_ij . G_ij + b_ij . G'_ij over in-cutoff neighbors,
 * a = c (ahat^T Ri - abar^T v), b = -c v_0 abar^T Ri, from abar = dE/dA and
 * ahat = dL/dA of the fitting side ([N][4][E], not scaled by c). chunk =
 * atoms per wavefront. The empty-slot term is the caller's.
 */
int matpl_dp_train_tangent(const double *pos, const double *lattice, const int32_t *atom_frame,
                           const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift, const int32_t *counts, const int32_t *type_map, const float *vec,
                           const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E,
                           int bf16, double rcut, double rcut_smooth, double scale, const float *resid, float *dA,
                           int32_t *nin, int32_t *status, void *stream);
int matpl_dp_train_embgrad(const double *pos, const double *lattice, const int32_t *atom_frame,
                           const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                           const int32_t *shift, const int32_t *counts, const int32_t *type_map, const float *vec,
                           const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E,
                           int bf16, double rcut, double rcut_smooth, double scale, const float *resid,
                           const float *abar, const float *ahat, float *grad, int chunk, int32_t *status,
                           void *stream);
int matpl_dp_train_grad_size(int E);
/*
 * Launch shape of the fused forward (kernel 0) or backward (kernel 1) kernel
 * for later calls in this process: the persistent grid's block cap
 * (default 4096), the weights staged in LDS (1), read from global memory (0)
 * or LDS when they fit (-1, default), and the XCD-aware block order (1) or
 * plain order (0, default). Results do not depend on the shape except through
 * the virial's FP64 summation order. No counterpart in the reference: it is
 * the knob the launch autotuner (matpl_amd/autotune.py) turns.
 */
int matpl_dp_fused_set_launch(int kernel, int max_blocks, int lds_weights, int xcd);
/* Wavefronts per block (2, 4 = default, or 8) of the net-mode fused forward
 * (kernel 0) / backward (kernel 1) for later calls; one atom per wavefront, so
 * this sets the atoms that share a block's staged weights. The autotuner's
 * second knob; no reference counterpart. */
int matpl_dp_fused_set_waves(int kernel, int waves);

/* force[i] = fcenter[i] + sum of fslot over the used entries of atom i's list
 * (counts [N][T]), in a fixed order (deterministic, no atomics). No reference
 * counterpart (the reference scatters with atomics, calculateForce.cu:92). */
int matpl_dp_fused_gather(const int32_t *counts, const float *fslot, const float *fcenter, float *force, int ntypes,
                          int mnl, int64_t natoms, void *stream);

/*
 * Fitting net of the fused path on the matrix cores: A [N][4][E] ->
 * Ei [N] and dA = dEtot/dA [N][4][E], per tile of 32 atoms of one centre
 * type (tiles [ntiles][3] = type, first atom, count). DR = A^T A[:, :16]
 * is formed on the fly (M2 = 16, E <= 32); hidden width <= 64, nlayers
 * (1..4) tanh layers, the first plain and the rest residual with resnet_dt,
 * then a linear output. Weights packed by matpl_amd/dp_fitting.py (bf16 or
 * fp32 per `bf16`). Replaces calculate_Ei / FittingNet.forward and its
 * autograd backward (src/model/dp_dp.py:304-415,
 * src/model/dp_embedding.py:96-163).
 */
int matpl_dp_fitting(const float *A, const int32_t *tiles, int ntiles, const void *w0f, const void *w0b,
                     const void *whf, const void *whb, const float *vec, int64_t type_stride_w0f,
                     int64_t type_stride_w0b, int64_t type_stride_wh, int64_t type_stride_vec, int E, int nlayers,
                     int bf16, float *Ei, float *dA, void *stream);
/* The same with tiles over positions of perm [Npos] (the atom at each
 * position; atoms grouped by type need not be contiguous, e.g. packed frames). */
int matpl_dp_fitting_perm(const float *A, const int32_t *tiles, const int32_t *perm, int ntiles, const void *w0f,
                          const void *w0b, const void *whf, const void *whb, const float *vec,
                          int64_t type_stride_w0f, int64_t type_stride_w0b, int64_t type_stride_wh,
                          int64_t type_stride_vec, int E, int nlayers, int bf16, float *Ei, float *dA, void *stream);
/*
 * The fitting nets' part of the force-loss double backward for training (the
 * reference's dE = autograd.grad(Etot, Ri, create_graph=True) and
 * loss.backward() through FittingNet, src/model/dp_dp.py:304-429,
 * src/PWMLFF/dp_mods/dp_trainer.py:274). Per atom, with At [N][4][E] the
 * tangent of A along the force residual and gE [N] = dL/dEi of the energy
 * loss (NULL: 0), the net runs forward with tangents and then the reverse
 * pass of L_i = gE_i E_i - <dE_i/dA_i, At_i> (the force loss reaches the
 * fitting side only through that term), giving Ahat [N][4][E] = dL/dA and,
 * per position of perm (tiles and weights as matpl_dp_fitting_perm), the
 * rows whose sums over each type's positions are the parameter gradients:
 * xw0 [Npos][2][16E] (DR, its tangent) and zw0 [Npos][2][64] (the layer-0
 * pre-activation adjoint and its tangent's adjoint): dW0 = sum xw0^T zw0,
 * db0 = sum zw0[., 0]; xh, zh [Npos][nlayers-1][2][64] likewise for each
 * hidden layer, rb [Npos][nlayers-1][64] the resnet_dt gradient terms,
 * xo [Npos][64] = gE x - x_tangent the output weight's; tz [Npos][nlayers]
 * [2][2][32] is scratch. Row buffers 16-byte aligned; hidden features past
 * the width are zero.
 */
int matpl_dp_fit_train(const float *A, const float *At, const float *gE, const int32_t *perm, const int32_t *tiles,
                       int ntiles, const void *w0f, const void *w0b, const void *whf, const void *whb,
                       const float *vec, int64_t type_stride_w0f, int64_t type_stride_w0b, int64_t type_stride_wh,
                       int64_t type_stride_vec, int E, int nlayers, int bf16, float *Ahat, float *xw0, float *zw0,
                       float *xh, float *zh, float *rb, float *xo, float *tz, void *stream);

/*
 * Y[t] = X_t^T Z_t [M][N] (float, written) for tall operands: the training
 * step's fitting-weight gradients, sums over each type's atoms of products of
 * per-atom rows (the reference leaves them to autograd's GEMMs,
 * src/model/dp_embedding.py:96-163 through loss.backward()). Type t's rows are
 * r in [2 start[t], 2 start[t+1]) (start: host array of ntypes + 1 entries,
 * ntypes <= 16); row r of X is at X + (r / 2) * ldx2 + (r % 2) * ldx1 (M
 * values), of Z likewise (N <= 64 values). work: matpl_tall_xtz_work(
 * max atoms of a type, ntypes, M, N) floats. Deterministic (per-chunk partial
 * products added in chunk order). No reference counterpart.
 */
int64_t matpl_tall_xtz_work(int64_t max_atoms, int ntypes, int M, int N);
int matpl_tall_xtz(const float *X, int64_t ldx2, int64_t ldx1, const float *Z, int64_t ldz2, int64_t ldz1,
                   const int64_t *start, int ntypes, int M, int N, float *Y, float *work, int64_t work_elems,
                   void *stream);

/*
 * Column sums of per-type row ranges over up to 8 buffers in one launch: the
 * training step's bias, resnet_dt and output-layer gradients (the reference
 * leaves them to autograd, src/model/dp_embedding.py:96-163, FittingNet, through loss.backward()).
 * For buffer s (ptrs, lds, cols: host arrays of nseg entries; ptrs holds
 * device addresses of float buffers) and type t,
 * out[t][off_s + j] = sum_{p = start[t]}^{start[t+1]-1} ptrs[s][p * lds[s] + j],
 * j < cols[s], off_s = cols[0] + ... + cols[s-1]; out is [ntypes][sum cols].
 * start: host array of ntypes + 1 entries (ntypes <= 16). work:
 * matpl_seg_rowsum_work(max positions of a type, ntypes, sum cols) floats.
 * Deterministic (fixed summation order). No reference counterpart.
 */
int64_t matpl_seg_rowsum_work(int64_t max_atoms, int ntypes, int total_cols);
int matpl_seg_rowsum(int nseg, const uint64_t *ptrs, const int64_t *lds, const int32_t *cols, const int64_t *start,
                     int ntypes, float *out, float *work, int64_t work_elems, void *stream);

/* ------------------------------------------------------------ DP compress */
/*
 * calculateCompress: G[s][i] = polynomial of degree coe_num-1 in f2[s] with
 * coefficients coef[s][i][0..coe_num-1], highest power first (coe_num 4 or
 * 6). f2 [n], coef [n][L][coe_num], G [n][L]. Replaces
 * CalculateCompressFuncs::forward (src/op/src/CalcOps.cpp:331-342) ->
 * compress_calc (src/op/kernel/calculateCompress.cu).
 */
int matpl_dp_compress_f32(const float *f2, const float *coef, int64_t n, int L, int coe_num, float *G, void *stream);
int matpl_dp_compress_f64(const double *f2, const double *coef, int64_t n, int L, int coe_num, double *G,
                          void *stream);
/*
 * Backward wrt f2: gf2[s] = sum_i gG[s][i] dG[s][i]/df2 (the reference forms
 * dG/df2 in compress_calc_grad, src/op/kernel/calculateCompress_grad.cu, and
 * the product and sum in torch, CalcOps.cpp:344-357).
 */
int matpl_dp_compress_grad_f32(const float *f2, const float *coef, const float *gG, int64_t n, int L, int coe_num,
                               float *gf2, void *stream);
int matpl_dp_compress_grad_f64(const double *f2, const double *coef, const double *gG, int64_t n, int L,
                               int coe_num, double *gf2, void *stream);

/*
 * calculateVirial without a workspace and without allocating device memory:
 * the same work, kernels and results as matpl_dp_virial_f32 (f64 = 0) /
 * matpl_dp_virial_f64 (f64 != 0), with the reduction's partial sums in a
 * static device array of the library (1.18 MB per GPU, held until the process
 * ends). B <= 64 frames, else MATPL_ERR_ARGUMENT. The array is looked up on the
 * current device, which must be the one the pointers and stream belong to.
 * Calls that overlap on one GPU share the array and race; serialize them (one
 * stream, or events). No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).
 */
int matpl_dp_virial_noalloc(int f64, const int32_t *nlist, const void *dE, const void *Rij, const void *Ri_d,
                            int64_t B, int64_t N, int64_t S, int64_t nghost, void *virial, void *atom_virial,
                            void *stream);

/* ------------------------------------------------------------ NEP operators */
/*
 * Neighbor search over replicated boxes (num_cell[f][3] replicas of the
 * original cell, then the minimum image of the expanded box, as the
 * reference's expand_box + gpu_find_neighbor_list). Frames are packed:
 * frame_start[nframes+1], atom_frame[n]; box/box_orig [nframes][9] (box =
 * expanded lattice), pos [n][3] Cartesian. mode 0: neighbor lists
 * NL_r[n][maxr], NL_a[n][maxa] (global atom index, -1 padded), Ri [n][max][3
 * or 4 with rij4] = [(r,) x, y, z], NN counts; *status = 1+atom if an atom
 * has more than maxr/maxa neighbors. mode 1: NN_r/NN_a[n] counts only.
 * mode 2: per-type counts NN_r/NN_a[n][ntypes] (type_map int64 [n]).
 * Replaces launch_calculate_neighbor / launch_calculate_maxneigh
 * (src/op/include/calculate_nepneighbor.h:1-42).
 */
int matpl_nep_neighbor(const double *box, const double *box_orig, const int64_t *num_cell,
                       const int64_t *frame_start, const int32_t *atom_frame, const int64_t *type_map,
                       const double *pos, double rc_r, double rc_a, int64_t maxr, int64_t maxa, int64_t n, int mode,
                       int ntypes, int rij4, int64_t *NN_r, int64_t *NL_r, int64_t *NN_a, int64_t *NL_a, double *Ri_r,
                       double *Ri_a, int32_t *status, void *stream);

/*
 * matpl_nep_neighbor's lists (mode 0) with float Ri rows: each value is the
 * FP64 one rounded once, so the rows equal a float cast of matpl_nep_neighbor's
 * (the reference's NEP model computes in Ri's dtype, src/model/nep_net.py:
 * 331-340, so a float model gets float rows; here the cast happens in the
 * search kernel). Same arguments as mode 0 otherwise.
 */
int matpl_nep_neighbor_ri32(const double *box, const double *box_orig, const int64_t *num_cell,
                            const int64_t *frame_start, const int32_t *atom_frame, const double *pos, double rc_r,
                            double rc_a, int64_t maxr, int64_t maxa, int64_t n, int rij4, int64_t *NN_r,
                            int64_t *NL_r, int64_t *NN_a, int64_t *NL_a, float *Ri_r, float *Ri_a, int32_t *status,
                            void *stream);

/*
 * NEP features of one family and their derivatives; f64 selects double,
 * else float. angular = 0: radial features q[n] = sum_j sum_k c[ti][tj][n][k]
 * f_k(r_ij); angular = 1: angular features from s_nlm sums (lmax l3 of the
 * 3-body, l4 > 0 adds the 4-body, l5 > 0 the 5-body term). c [T][T][nm][nb],
 * Ri [N][S][4] = [r, x, y, z], NL [N][S] (-1 padded), atom_map [N];
 * features written to feats[i*ld + col0 + ...].
 *  mode 0: feats (+ s_save [N][nm*24] for angular);
 *  mode 1: gRi [N][S][4] = dL/dRi and gc += dL/dc for gfeat = dL/dfeats;
 *  mode 2: second order for gg = upstream grad of gRi: ggfeat [N][ld] and gc.
 * Replaces launch_calculate_nepfeat / _grad / _secondgradout(_c2) and
 * launch_calculate_nepmbfeat / _grad / _secondgradout(_c3)
 * (src/op/include/calculate_nepfeat.h:46-175).
 */
int matpl_nep_feat(int f64, int angular, int mode, const void *c, const void *Ri, const int64_t *NL,
                   const int64_t *atom_map, int64_t N, int S, int T, int nm, int nb, int ld, int col0, int l3, int l4,
                   int l5, double rc, void *feats, void *s_save, const void *gfeat, const void *gg, void *gRi,
                   void *gc, void *ggfeat, void *stream);

/*
 * Neighbor scatter F[NL[i][k]] += dE[i][k] . Ri_d[i][k] (a 4x3 contraction)
 * into F [N][3], which holds the centre terms -sum_k dE[i][k] . Ri_d[i][k]
 * on entry. dE [N][S][4], Ri_d [N][S][4][3]. Replaces
 * launch_calculate_nepforce (src/op/include/calculate_nepfeat.h:1).
 */
int matpl_nep_force(int f64, const int64_t *NL, const void *dE, const void *Ri_d, void *F, int64_t N, int S,
                    void *stream);

/*
 * Per-frame virial [nframes][9] and per-atom virial [N][9] from the same
 * pair terms and Rij [N][S][3]; frames are atom ranges frame_start[f] ..
 * frame_start[f+1]. Replaces launch_calculate_nepvirial
 * (src/op/include/calculate_nepfeat.h:21).
 */
int matpl_nep_virial(int f64, const int64_t *NL, const void *dE, const void *Rij, const void *Ri_d,
                     const int64_t *frame_start, int64_t nframes, int64_t N, int S, void *virial, void *atom_virial,
                     void *stream);

/*
 * Centre term of the NEP force, F[i] = -sum_{k,w} dE[i][k][w] Ri_d[i][k][w][:]
 * over all S slots, F [N][3]. No counterpart in the reference's operator
 * library: it replaces the torch.matmul that the reference's NEP model uses to
 * form this term before calculateNepForce (src/model/nep_net.py:608).
 */
int matpl_nep_centre_force(int f64, const void *dE, const void *Ri_d, int64_t N, int S, void *F, void *stream);

/*
 * Per-frame virial [nframes][9] alone: the first output of matpl_nep_virial
 * without its per-atom rows (indexed by neighbor), for callers that do not use
 * atom_virial (the reference's NEP model returns only the frame virial,
 * src/model/nep_net.py:438, 613-622). Frame f holds atoms
 * frame_start[f]..frame_start[f+1] (int64 [nframes+1]); each list entry counts
 * toward its centre atom's frame (lists never leave a frame, where this equals
 * the sum of the per-atom rows). rows [N][9] is a workspace that receives each
 * atom's centre-row sum; the frame sums are formed from it without atomics
 * (deterministic). Same pair terms as launch_calculate_nepvirial
 * (src/op/include/calculate_nepfeat.h:21).
 */
int matpl_nep_frame_virial(int f64, const int64_t *NL, const void *dE, const void *Rij, const void *Ri_d,
                           const int64_t *frame_start, int64_t nframes, int64_t N, int S, void *rows, void *virial,
                           void *stream);

/*
 * Backward of the whole force (centre + neighbor terms) with respect to dE:
 * g[i][k] = Ri_d[i][k] . (gF[NL[i][k]] - gF[i]), g [N][S][4] from gF [N][3].
 * Replaces launch_calculate_nepforce_grad (src/op/include/calculate_nepfeat.h:11).
 */
int matpl_nep_force_grad(int f64, const int64_t *NL, const void *Ri_d, const void *gF, void *g, int64_t N, int S,
                         void *stream);

/*
 * Backward of matpl_nep_virial with respect to dE, with gV [nframes][9]
 * applied to each atom's own frame (atom_frame [N]). Replaces
 * launch_calculate_nepvirial_grad (src/op/include/calculate_nepfeat.h:35).
 */
int matpl_nep_virial_grad(int f64, const int64_t *NL, const void *Rij, const void *Ri_d, const void *gV,
                          const int32_t *atom_frame, void *g, int64_t N, int S, void *stream);

/*
 * The NEP model's force and frame virial of one neighbor list in one pass,
 * from dE [N][S][4] and Ri [N][S][4] = [r, x, y, z] (both 16-byte aligned):
 * the position derivative Ri_d (row 0 = (x, y, z)/r, rows 1-3 = identity,
 * zero where |r| <= 1e-5) is formed in registers instead of read. Fuses the
 * model's sequence ri_d (src/model/nep_net.py:440-468), the centre term
 * (nep_net.py:608), calculateNepForce (src/op/src/CalcOps.cpp:75-158) and the
 * frame sum of calculateNepVirial (CalcOps.cpp:239-328, nep_net.py:613-622):
 * F [N][3] = the whole force (centre + neighbor terms; written, not
 * accumulated), virial [nframes][9] per frame. Frame f holds atoms
 * frame_start[f]..frame_start[f+1]; atom_frame [N]; every list entry must lie
 * in its atom's frame; max_frame >= the largest frame's atom count. work:
 * matpl_nep_force_virial_ri_work(f64, N, max_frame) elements of the dtype
 * (per-atom virial rows and per-block LDS force copies). Entries with
 * NL < 0 contribute nothing.
 */
int64_t matpl_nep_force_virial_ri_work(int f64, int64_t N, int64_t max_frame);
int matpl_nep_force_virial_ri(int f64, const int64_t *NL, const void *dE, const void *Ri, const int64_t *frame_start,
                              const int32_t *atom_frame, int64_t nframes, int64_t N, int S, int64_t max_frame,
                              void *F, void *virial, void *work, int64_t work_elems, void *stream);

/*
 * Backward of matpl_nep_force_virial_ri with respect to dE: g [N][S][4]
 * (written) from gF [N][3] and gV [nframes][9] (either may be NULL for zero;
 * atom_frame is needed with gV): with q = gF[j] - gF[i] + gV[frame(i)] Rij,
 * g = [q . (x, y, z)/r, q] on used entries with |r| > 1e-5, else 0. The sum of
 * matpl_nep_force_grad and matpl_nep_virial_grad without Ri_d.
 */
int matpl_nep_force_virial_ri_grad(int f64, const int64_t *NL, const void *Ri, const void *gF, const void *gV,
                                   const int32_t *atom_frame, int64_t N, int S, void *g, void *stream);

/*
 * NEP fitting nets (one hidden layer of H tanh units, linear output; the
 * reference's FittingNet with network_size [H, 1], src/model/nep_fitting.py:
 * 40-163, run per atom type at src/model/nep_net.py:470-501) for all types in
 * one launch. f64: double, else float, for every floating argument.
 * q [N][D] features, qs [D] q_scaler; perm [N] the atoms grouped by type and
 * tiles [ntiles][3] (type, first position in perm, count <= tile_atoms), the
 * tiles of one type consecutive and the types in increasing order; tile_atoms
 * must be matpl_nep_fit_tile_atoms(f64, D, H) (-1: D, H too large).
 * W0 [T][D][H], b0 [T][H], w1 [T][H] (the output layer's weight column),
 * b1 [T].
 * Forward: Ei [N] and, when g is not NULL, g [N][D] = dEi/dq.
 * Backward: given gE [N] = dL/dEi and gg [N][D] = dL/dg (either may be NULL
 * for zero), dq [N][D] = dL/dq (may be NULL) and the parameter gradients
 * dW0/db0/dw1/db1 (written, not accumulated) of L through Ei and g, i.e. the
 * net's first and second derivatives (the force loss reaches the weights
 * through g). part: workspace of ntiles * matpl_nep_fit_part_width(D, H)
 * elements; tstart [T+1]: type t's tiles are [tstart[t], tstart[t+1]).
 * Results do not depend on scheduling (per-tile sums added in a fixed order).
 */
int matpl_nep_fit_tile_atoms(int f64, int D, int H);
int matpl_nep_fit_part_width(int D, int H);
int matpl_nep_fit_forward(int f64, const void *q, const void *qs, const int32_t *perm, const int32_t *tiles,
                          int ntiles, int tile_atoms, const void *W0, const void *b0, const void *w1, const void *b1,
                          int D, int H, void *Ei, void *g, void *stream);
int matpl_nep_fit_backward(int f64, const void *q, const void *qs, const int32_t *perm, const int32_t *tiles,
                           int ntiles, int tile_atoms, const int32_t *tstart, int ntypes, const void *W0,
                           const void *b0, const void *w1, const void *b1, int D, int H, const void *gE,
                           const void *gg, void *dq, void *part, void *dW0, void *db0, void *dw1, void *db1,
                           void *stream);

/* ------------------------------------------------------------ MD helpers */
/*
 * Velocity Verlet for MD steps driven by the fused path (units A, fs, eV,
 * amu): v += dt/2 F/m, and with drift != 0 also x += dt v. v, x [n][3]
 * float64, F [n][3] float, inv_mass [n] float64 (1/amu). No reference
 * counterpart: in the reference the MD integrator is LAMMPS's, which calls
 * the model through PairMATPL::compute (src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:864).
 */
int matpl_md_kick(double *v, double *x, const float *F, const double *inv_mass, double dt, int64_t n, int drift,
                  void *stream);
/* out[i] = in[i] . M for row vectors (M host [9]); Cartesian <-> fractional. */
int matpl_md_affine(const double *in, double *out, const double *M_host, int64_t n, void *stream);

/* ------------------------------------------------ tall-by-small products */
/*
 * y [K][b] = x [K][a] . W [a][b], FP32, 1 <= a, b <= 64: x row-major with row
 * stride ldx >= a, W element (j, c) at W[j * w_row + c * w_col], y contiguous.
 * lowp != 0: operands and result rounded to bfloat16 (products and sums in FP32).
 * Replaces the layer products torch.matmul(x, weight) of EmbeddingNet /
 * FittingNet (src/model/dp_embedding.py:73-94, 140-152) in the batched DP
 * training step, where the reference leaves them to the BLAS library.
 */
int matpl_tall_mm(const float *x, int64_t ldx, const float *W, int64_t w_row, int64_t w_col, float *y, int64_t K,
                  int a, int b, int lowp, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MATPL_HIP_H */
