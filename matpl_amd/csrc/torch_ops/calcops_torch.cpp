// TORCH_LIBRARY registration of the ten CalcOps_cuda operators on top of
// libmatpl_hip.so, built as libCalcOps_bind.so: the name and namespace the
// reference's models load (torch.ops.load_library(<src>/op/build/lib/
// libCalcOps_bind.so), src/model/dp_dp.py:13-20, src/model/nep_net.py:17-24)
// and that TorchScript models and the LAMMPS pair style
// (src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp) find through OP_LIB_PATH.
//
// Same names, schemas (inferred from the C++ signatures, as the reference's
// m.def(name, fn) does) and autograd edges as the reference's bindings
// (src/op/src/CalcOps_bind.cpp:6-17, src/op/include/CalcOps.h:35-318):
//   calculateForce   -> [force [B][N+nghost][3]];   backward to dE only
//   calculateVirial  -> [virial [B][9], atom_virial [B][N+nghost][9]]; backward to dE
//   calculateCompress-> [G [n][L]];                   backward to f2 only
//   the seven NEP operators: nep_ops_torch.cpp.
// Work is enqueued on torch's current HIP stream; errors raise (TORCH_CHECK)
// with the library's message.
#include <torch/library.h>
#include <torch/torch.h>
#include <c10/hip/HIPStream.h>

#include "../../../include/matpl_hip.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void *stream() { return (void *)c10::hip::getCurrentHIPStream().stream(); }

void check(int rc, const char *what) { TORCH_CHECK(rc == 0, what, ": ", matpl_last_error()); }

int64_t ghosts(const Tensor &n) { return n.defined() && n.numel() ? n.item<int64_t>() : 0; }

Tensor aligned(const Tensor &t) {
    Tensor c = t.contiguous();
    if (reinterpret_cast<uintptr_t>(c.data_ptr()) % 16) c = c.clone();
    return c;
}

void check_float(const Tensor &t, const char *op) {
    TORCH_CHECK(t.is_cuda(), op, ": tensors must be on the GPU");
    TORCH_CHECK(t.scalar_type() == torch::kFloat || t.scalar_type() == torch::kDouble, op,
                ": unsupported dtype (float32 / float64)");
}

// list_neigh [B, N, 1?, S] -> (B, N, S)
void dims(const Tensor &nl, int64_t &B, int64_t &N, int64_t &S) {
    B = nl.size(0);
    N = nl.size(1);
    S = B * N ? nl.numel() / (B * N) : 0;
}

class ForceFn : public torch::autograd::Function<ForceFn> {
  public:
    static Tensor forward(AutogradContext *ctx, Tensor nl, Tensor dE, Tensor Ri_d, Tensor F, Tensor nghost_t) {
        check_float(dE, "calculateForce");
        int64_t B, N, S;
        dims(nl, B, N, S);
        const int64_t ng = ghosts(nghost_t);
        Tensor nli = nl.to(torch::kInt).contiguous();
        Tensor e = aligned(dE), d = aligned(Ri_d.to(dE.scalar_type()));
        Tensor f = F.to(dE.scalar_type()).contiguous();
        Tensor out = torch::empty({B, N + ng, 3}, dE.options());
        const bool f64 = dE.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_force_f64(nli.data_ptr<int32_t>(), e.data_ptr<double>(), d.data_ptr<double>(),
                                       f.data_ptr<double>(), B, N, S, ng, out.data_ptr<double>(), stream())
                  : matpl_dp_force_f32(nli.data_ptr<int32_t>(), e.data_ptr<float>(), d.data_ptr<float>(),
                                       f.data_ptr<float>(), B, N, S, ng, out.data_ptr<float>(), stream()),
              "calculateForce");
        ctx->save_for_backward({nli, d});
        ctx->saved_data["ng"] = ng;
        ctx->saved_data["shape"] = dE.sizes().vec();
        return out;
    }

    static variable_list backward(AutogradContext *ctx, variable_list grads) {
        auto sv = ctx->get_saved_variables();
        Tensor nli = sv[0], d = sv[1];
        int64_t B, N, S;
        dims(nli, B, N, S);
        const int64_t ng = ctx->saved_data["ng"].toInt();
        Tensor gF = grads[0].to(d.scalar_type()).contiguous();
        Tensor g = torch::empty(ctx->saved_data["shape"].toIntVector(), d.options());
        const bool f64 = d.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_force_grad_f64(nli.data_ptr<int32_t>(), d.data_ptr<double>(), gF.data_ptr<double>(), B,
                                            N, S, ng, g.data_ptr<double>(), stream())
                  : matpl_dp_force_grad_f32(nli.data_ptr<int32_t>(), d.data_ptr<float>(), gF.data_ptr<float>(), B,
                                            N, S, ng, g.data_ptr<float>(), stream()),
              "calculateForce backward");
        return {Tensor(), g, Tensor(), Tensor(), Tensor()};
    }
};

class VirialFn : public torch::autograd::Function<VirialFn> {
  public:
    static variable_list forward(AutogradContext *ctx, Tensor nl, Tensor dE, Tensor Rij, Tensor Ri_d,
                                 Tensor nghost_t) {
        check_float(dE, "calculateVirial");
        int64_t B, N, S;
        dims(nl, B, N, S);
        const int64_t ng = ghosts(nghost_t);
        Tensor nli = nl.to(torch::kInt).contiguous();
        Tensor e = aligned(dE), d = aligned(Ri_d.to(dE.scalar_type()));
        Tensor r = Rij.to(dE.scalar_type()).contiguous();
        Tensor v = torch::empty({B, 9}, dE.options());
        Tensor av = torch::empty({B, N + ng, 9}, dE.options());
        const int64_t ws_elems = std::max<int64_t>(matpl_dp_virial_workspace_elems(B, N, ng), 1);
        Tensor ws = torch::empty({ws_elems}, dE.options());
        const bool f64 = dE.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_virial_f64(nli.data_ptr<int32_t>(), e.data_ptr<double>(), r.data_ptr<double>(),
                                        d.data_ptr<double>(), B, N, S, ng, v.data_ptr<double>(),
                                        av.data_ptr<double>(), ws.data_ptr<double>(), ws_elems, stream())
                  : matpl_dp_virial_f32(nli.data_ptr<int32_t>(), e.data_ptr<float>(), r.data_ptr<float>(),
                                        d.data_ptr<float>(), B, N, S, ng, v.data_ptr<float>(), av.data_ptr<float>(),
                                        ws.data_ptr<float>(), ws_elems, stream()),
              "calculateVirial");
        ctx->save_for_backward({nli, r, d});
        ctx->saved_data["shape"] = dE.sizes().vec();
        ctx->mark_non_differentiable({av});
        return {v, av};
    }

    static variable_list backward(AutogradContext *ctx, variable_list grads) {
        if (!grads[0].defined()) return {Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
        auto sv = ctx->get_saved_variables();
        Tensor nli = sv[0], r = sv[1], d = sv[2];
        int64_t B, N, S;
        dims(nli, B, N, S);
        Tensor gV = grads[0].to(d.scalar_type()).contiguous();
        Tensor g = torch::empty(ctx->saved_data["shape"].toIntVector(), d.options());
        const bool f64 = d.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_virial_grad_f64(nli.data_ptr<int32_t>(), r.data_ptr<double>(), d.data_ptr<double>(),
                                             gV.data_ptr<double>(), B, N, S, g.data_ptr<double>(), stream())
                  : matpl_dp_virial_grad_f32(nli.data_ptr<int32_t>(), r.data_ptr<float>(), d.data_ptr<float>(),
                                             gV.data_ptr<float>(), B, N, S, g.data_ptr<float>(), stream()),
              "calculateVirial backward");
        return {Tensor(), g, Tensor(), Tensor(), Tensor()};
    }
};

class CompressFn : public torch::autograd::Function<CompressFn> {
  public:
    static Tensor forward(AutogradContext *ctx, Tensor f2, Tensor coef) {
        check_float(f2, "calculateCompress");
        TORCH_CHECK(coef.dim() == 3 && (coef.size(2) == 4 || coef.size(2) == 6),
                    "calculateCompress: coefficient must be [n, L, 4|6]");
        TORCH_CHECK(f2.numel() == coef.size(0), "calculateCompress: f2 and coefficient rows differ");
        Tensor x = f2.contiguous(), c = coef.to(f2.scalar_type()).contiguous();
        const int64_t n = c.size(0);
        const int L = (int)c.size(1), K = (int)c.size(2);
        Tensor G = torch::empty({n, L}, f2.options());
        const bool f64 = f2.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_compress_f64(x.data_ptr<double>(), c.data_ptr<double>(), n, L, K, G.data_ptr<double>(),
                                          stream())
                  : matpl_dp_compress_f32(x.data_ptr<float>(), c.data_ptr<float>(), n, L, K, G.data_ptr<float>(),
                                          stream()),
              "calculateCompress");
        ctx->save_for_backward({x, c});
        ctx->saved_data["shape"] = f2.sizes().vec();
        return G;
    }

    static variable_list backward(AutogradContext *ctx, variable_list grads) {
        auto sv = ctx->get_saved_variables();
        Tensor x = sv[0], c = sv[1];
        const int64_t n = c.size(0);
        const int L = (int)c.size(1), K = (int)c.size(2);
        Tensor gG = grads[0].to(x.scalar_type()).contiguous();
        Tensor g = torch::empty({n}, x.options());
        const bool f64 = x.scalar_type() == torch::kDouble;
        check(f64 ? matpl_dp_compress_grad_f64(x.data_ptr<double>(), c.data_ptr<double>(), gG.data_ptr<double>(), n,
                                               L, K, g.data_ptr<double>(), stream())
                  : matpl_dp_compress_grad_f32(x.data_ptr<float>(), c.data_ptr<float>(), gG.data_ptr<float>(), n, L,
                                               K, g.data_ptr<float>(), stream()),
              "calculateCompress backward");
        return {g.reshape(ctx->saved_data["shape"].toIntVector()), Tensor()};
    }
};

std::vector<Tensor> calculateForce(Tensor list_neigh, Tensor dE, Tensor Ri_d, Tensor F, Tensor nghost_tensor) {
    return {ForceFn::apply(list_neigh, dE, Ri_d, F, nghost_tensor)};
}

std::vector<Tensor> calculateVirial(Tensor list_neigh, Tensor dE, Tensor Rij, Tensor Ri_d, Tensor nghost_tensor) {
    auto r = VirialFn::apply(list_neigh, dE, Rij, Ri_d, nghost_tensor);
    return {r[0], r[1]};
}

std::vector<Tensor> calculateCompress(Tensor f2, Tensor coefficient) { return {CompressFn::apply(f2, coefficient)}; }

// The build's own environment-matrix op (DP.calculate_Ri, src/model/dp_dp.py:217-302), for
// scripted models: [Ri [B][N][S][4], Ri_d [B][N][S][4][3]] from ImageDR [B][N][S][4]. No autograd
// edge: models treat Ri as the leaf they differentiate the energy by, Ri_d is analytic.
std::vector<Tensor> env_mat(Tensor image_dr, Tensor type_map, Tensor davg_full, Tensor dstd_full, double rmin,
                            double rmax) {
    check_float(image_dr, "env_mat");
    Tensor idr = image_dr.dim() == 3 ? image_dr.unsqueeze(0).contiguous() : image_dr.contiguous();
    TORCH_CHECK(idr.dim() == 4 && idr.size(3) == 4, "env_mat: ImageDR must be [B, N, S, 4]");
    const int64_t B = idr.size(0), N = idr.size(1), S = idr.size(2);
    const int T = (int)davg_full.size(0);
    TORCH_CHECK(davg_full.numel() == (int64_t)T * S * 4 && dstd_full.numel() == (int64_t)T * S * 4,
                "env_mat: davg/dstd must be [T, S*4]");
    Tensor tm = type_map.to(torch::kInt).contiguous();
    Tensor da = davg_full.to(idr.scalar_type()).contiguous(), ds = dstd_full.to(idr.scalar_type()).contiguous();
    Tensor Ri = torch::empty({B, N, S, 4}, idr.options());
    Tensor Rid = torch::empty({B, N, S, 4, 3}, idr.options());
    const bool f64 = idr.scalar_type() == torch::kDouble;
    check(f64 ? matpl_dp_env_mat_f64(idr.data_ptr<double>(), tm.data_ptr<int32_t>(), da.data_ptr<double>(),
                                     ds.data_ptr<double>(), B, N, S, T, rmin, rmax, Ri.data_ptr<double>(),
                                     Rid.data_ptr<double>(), stream())
              : matpl_dp_env_mat_f32(idr.data_ptr<float>(), tm.data_ptr<int32_t>(), da.data_ptr<float>(),
                                     ds.data_ptr<float>(), B, N, S, T, rmin, rmax, Ri.data_ptr<float>(),
                                     Rid.data_ptr<float>(), stream()),
          "env_mat");
    return {Ri, Rid};
}

}  // namespace

TORCH_LIBRARY(matpl_amd, m) {
    m.def("env_mat(Tensor image_dr, Tensor type_map, Tensor davg_full, Tensor dstd_full, float rmin, float rmax) "
          "-> Tensor[]", env_mat);
}

namespace matpl_torch {
std::vector<Tensor> calculateNepFeat(Tensor coeff2, Tensor d12_radial, Tensor NL_radial, Tensor atom_map,
                                     Tensor feats, double rcut_radial, int64_t feat_multi_nums);
std::vector<Tensor> calculateNepMbFeat(Tensor coeff3, Tensor d12_angluar, Tensor NL_angluar, Tensor atom_map,
                                       Tensor feats, int64_t feat_2b_num, int64_t lmax_3, int64_t lmax_4,
                                       int64_t lmax_5, double rcut_angluar);
std::vector<Tensor> calculateNepForce(Tensor list_neigh, Tensor dE, Tensor Ri_d, Tensor F);
std::vector<Tensor> calculateNepVirial(Tensor list_neigh, Tensor dE, Tensor Rij, Tensor Ri_d, Tensor num_atom);
std::vector<Tensor> calculate_maxneigh(const Tensor &num_atoms, const Tensor &box, const Tensor &box_orig,
                                       const Tensor &num_cell, const Tensor &position, double cutoff_2b,
                                       double cutoff_3b, int64_t atom_type_num, const Tensor &atom_type_map,
                                       bool with_type);
std::vector<Tensor> calculate_neighbor(const Tensor &num_atoms, const Tensor &atom_type_map, const Tensor &atom_types,
                                       const Tensor &box, const Tensor &box_orig, const Tensor &num_cell,
                                       const Tensor &position, double cutoff_2b, double cutoff_3b,
                                       int64_t max_NN_radial, int64_t max_NN_angular, bool with_rij);
std::vector<Tensor> calculate_descriptor(const Tensor &weight_radial, const Tensor &weight_angular,
                                         const Tensor &Ri_radial, const Tensor &NL_radial,
                                         const Tensor &atom_type_map, double cutoff_radial, double cutoff_angular,
                                         int64_t max_NN_radial, int64_t lmax_3, int64_t lmax_4, int64_t lmax_5);
}  // namespace matpl_torch

TORCH_LIBRARY(CalcOps_cuda, m) {
    m.def("calculateForce", calculateForce);
    m.def("calculateVirial", calculateVirial);
    m.def("calculateCompress", calculateCompress);
    m.def("calculateNepFeat", matpl_torch::calculateNepFeat);
    m.def("calculateNepMbFeat", matpl_torch::calculateNepMbFeat);
    m.def("calculateNepForce", matpl_torch::calculateNepForce);
    m.def("calculateNepVirial", matpl_torch::calculateNepVirial);
    m.def("calculate_maxneigh", matpl_torch::calculate_maxneigh);
    m.def("calculate_neighbor", matpl_torch::calculate_neighbor);
    m.def("calculate_descriptor", matpl_torch::calculate_descriptor);
}
