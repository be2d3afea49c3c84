// The seven NEP operators of CalcOps_cuda as C++ autograd functions over
// libmatpl_hip.so, registered with the DP operators in calcops_torch.cpp so
// that one library holds all ten names of src/op/src/CalcOps_bind.cpp:6-17.
//
// Autograd edges follow the reference (src/op/src/CalcOps.cpp:385-1008):
//   calculateNepFeat / calculateNepMbFeat: features; backward to the
//     coefficients and to d12 through a gradient function whose own backward
//     (second order, for force-loss training) reaches the upstream feature
//     gradient and the coefficients;
//   calculateNepForce: backward to dE; calculateNepVirial: backward to dE from
//     the frame virial (each atom's own frame, where the reference applies
//     frame 0's gradient to every frame, calculateNepVirial.cu:110-157);
//   calculate_maxneigh / calculate_neighbor / calculate_descriptor: no edges
//     except calculate_descriptor's features (through the feature functions).
// float32 and float64 are both accepted (the reference casts to double); the
// feature operators return new tensors and leave `feats` untouched, and
// calculateNepForce returns a new tensor instead of writing into F.
#include <torch/library.h>
#include <torch/torch.h>
#include <c10/hip/HIPStream.h>

#include "../../../include/matpl_hip.h"

namespace matpl_torch {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

namespace {

void *stream() { return (void *)c10::hip::getCurrentHIPStream().stream(); }

void check(int rc, const char *what) { TORCH_CHECK(rc == 0, what, ": ", matpl_last_error()); }

void *ptr(const Tensor &t) { return t.defined() ? t.data_ptr() : nullptr; }

int f64_of(const Tensor &t, const char *op) {
    TORCH_CHECK(t.is_cuda(), op, ": tensors must be on the GPU");
    TORCH_CHECK(t.scalar_type() == torch::kFloat || t.scalar_type() == torch::kDouble, op,
                ": unsupported dtype (float32 / float64)");
    return t.scalar_type() == torch::kDouble ? 1 : 0;
}

int64_t n_angular(int64_t nm, int64_t l3, int64_t l4, int64_t l5) {
    return nm * l3 + (l4 > 0 ? nm : 0) + (l5 > 0 ? nm : 0);
}

// Checked, contiguous operands of a feature call: coefficients [T][T][n][k] in Ri's dtype,
// Ri [N][S][4], NL [N][S] int64 within the batch, atom_map [N] int64 below T.
struct FeatIn {
    Tensor c, Ri, NL, am;
};

FeatIn prep(const Tensor &c, const Tensor &Ri, const Tensor &NL, const Tensor &am, const char *op) {
    FeatIn in;
    in.Ri = Ri.contiguous();
    f64_of(in.Ri, op);
    in.NL = NL.to(torch::kLong).contiguous();
    in.am = am.to(torch::kLong).contiguous();
    TORCH_CHECK(c.dim() == 4 && c.size(0) == c.size(1), op, ": coefficients must be [T, T, n, k]");
    TORCH_CHECK(in.NL.dim() == 2 && in.am.numel() == in.NL.size(0), op,
                ": NL must be [N, S] with one atom_map entry per atom");
    TORCH_CHECK(in.Ri.numel() == in.NL.numel() * 4, op, ": d12 must be [N, S, 4]");
    if (in.NL.numel())
        TORCH_CHECK(in.NL.max().item<int64_t>() < in.NL.size(0), op, ": neighbor index beyond the batch");
    if (in.am.numel()) TORCH_CHECK(in.am.max().item<int64_t>() < c.size(0), op, ": atom type beyond the table");
    in.c = c.to(in.Ri.scalar_type()).contiguous();
    return in;
}

struct Ang {
    int64_t l3 = 0, l4 = 0, l5 = 0;
};

void feat_call(int mode, bool angular, const FeatIn &in, double rc, Ang l, const Tensor &feats,
               const Tensor &s_save, const Tensor &gfeat, const Tensor &gg, const Tensor &gRi, const Tensor &gc,
               const Tensor &ggfeat) {
    const int64_t N = in.NL.size(0);
    const int S = (int)in.NL.size(1);
    const int T = (int)in.c.size(0), nm = (int)in.c.size(2), nb = (int)in.c.size(3);
    const int ld = (int)(angular ? n_angular(nm, l.l3, l.l4, l.l5) : nm);
    check(matpl_nep_feat(f64_of(in.Ri, "nep feat"), angular ? 1 : 0, mode, in.c.data_ptr(), in.Ri.data_ptr(),
                         in.NL.data_ptr<int64_t>(), in.am.data_ptr<int64_t>(), N, S, T, nm, nb, ld, 0, (int)l.l3,
                         (int)l.l4, (int)l.l5, rc, ptr(feats), ptr(s_save), ptr(gfeat), ptr(gg), ptr(gRi), ptr(gc),
                         ptr(ggfeat), stream()),
          angular ? "calculateNepMbFeat" : "calculateNepFeat");
}

// ------------------------------------------------------------------ radial
class RadialGrad : public torch::autograd::Function<RadialGrad> {
  public:
    static variable_list forward(AutogradContext *ctx, Tensor g, Tensor c2, Tensor Ri, Tensor NL, Tensor am,
                                 double rc) {
        FeatIn in{c2, Ri, NL, am};
        Tensor gRi = torch::empty({NL.size(0), NL.size(1), 4}, Ri.options());
        Tensor gc2 = torch::zeros_like(c2);
        Tensor gin = g.contiguous();
        feat_call(1, false, in, rc, Ang{}, Tensor(), Tensor(), gin, Tensor(), gRi, gc2, Tensor());
        ctx->save_for_backward({gin, c2, Ri, NL, am});
        ctx->saved_data["rc"] = rc;
        return {gc2, gRi};
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        auto sv = ctx->get_saved_variables();
        const Tensor &ggRi = gout[1];
        if (!ggRi.defined()) return {Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
        FeatIn in{sv[1], sv[2], sv[3], sv[4]};
        Tensor ggfeat = torch::empty_like(sv[0]);
        Tensor gc2 = torch::zeros_like(sv[1]);
        feat_call(2, false, in, ctx->saved_data["rc"].toDouble(), Ang{}, Tensor(), Tensor(), sv[0],
                  ggRi.contiguous(), Tensor(), gc2, ggfeat);
        return {ggfeat, gc2, Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

class Radial : public torch::autograd::Function<Radial> {
  public:
    static Tensor forward(AutogradContext *ctx, Tensor c2, Tensor Ri, Tensor NL, Tensor am, double rc) {
        FeatIn in = prep(c2, Ri, NL, am, "calculateNepFeat");
        Tensor feats = torch::empty({in.NL.size(0), in.c.size(2)}, in.Ri.options());
        feat_call(0, false, in, rc, Ang{}, feats, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor());
        ctx->save_for_backward({in.c, in.Ri, in.NL, in.am});
        ctx->saved_data["rc"] = rc;
        return feats;
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        auto sv = ctx->get_saved_variables();
        auto r = RadialGrad::apply(gout[0].contiguous(), sv[0], sv[1], sv[2], sv[3], ctx->saved_data["rc"].toDouble());
        return {r[0], r[1], Tensor(), Tensor(), Tensor()};
    }
};

// ------------------------------------------------------------------ angular
class AngularGrad : public torch::autograd::Function<AngularGrad> {
  public:
    static variable_list forward(AutogradContext *ctx, Tensor g, Tensor c3, Tensor Ri, Tensor NL, Tensor am,
                                 Tensor s_save, double rc, int64_t l3, int64_t l4, int64_t l5) {
        FeatIn in{c3, Ri, NL, am};
        Tensor gRi = torch::empty({NL.size(0), NL.size(1), 4}, Ri.options());
        Tensor gc3 = torch::zeros_like(c3);
        Tensor gin = g.contiguous();
        feat_call(1, true, in, rc, Ang{l3, l4, l5}, Tensor(), s_save, gin, Tensor(), gRi, gc3, Tensor());
        ctx->save_for_backward({gin, c3, Ri, NL, am, s_save});
        ctx->saved_data["rc"] = rc;
        ctx->saved_data["l3"] = l3;
        ctx->saved_data["l4"] = l4;
        ctx->saved_data["l5"] = l5;
        return {gc3, gRi};
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        variable_list none(10);
        auto sv = ctx->get_saved_variables();
        const Tensor &ggRi = gout[1];
        if (!ggRi.defined()) return none;
        FeatIn in{sv[1], sv[2], sv[3], sv[4]};
        Ang l{ctx->saved_data["l3"].toInt(), ctx->saved_data["l4"].toInt(), ctx->saved_data["l5"].toInt()};
        Tensor ggfeat = torch::empty_like(sv[0]);
        Tensor gc3 = torch::zeros_like(sv[1]);
        feat_call(2, true, in, ctx->saved_data["rc"].toDouble(), l, Tensor(), sv[5], sv[0], ggRi.contiguous(),
                  Tensor(), gc3, ggfeat);
        none[0] = ggfeat;
        none[1] = gc3;
        return none;
    }
};

class Angular : public torch::autograd::Function<Angular> {
  public:
    static Tensor forward(AutogradContext *ctx, Tensor c3, Tensor Ri, Tensor NL, Tensor am, double rc, int64_t l3,
                          int64_t l4, int64_t l5) {
        FeatIn in = prep(c3, Ri, NL, am, "calculateNepMbFeat");
        const int64_t N = in.NL.size(0), nm = in.c.size(2);
        Tensor feats = torch::empty({N, n_angular(nm, l3, l4, l5)}, in.Ri.options());
        Tensor s_save = torch::empty({N, nm * 24}, in.Ri.options());
        feat_call(0, true, in, rc, Ang{l3, l4, l5}, feats, s_save, Tensor(), Tensor(), Tensor(), Tensor(), Tensor());
        ctx->save_for_backward({in.c, in.Ri, in.NL, in.am, s_save});
        ctx->saved_data["rc"] = rc;
        ctx->saved_data["l3"] = l3;
        ctx->saved_data["l4"] = l4;
        ctx->saved_data["l5"] = l5;
        return feats;
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        auto sv = ctx->get_saved_variables();
        auto r = AngularGrad::apply(gout[0].contiguous(), sv[0], sv[1], sv[2], sv[3], sv[4],
                                    ctx->saved_data["rc"].toDouble(), ctx->saved_data["l3"].toInt(),
                                    ctx->saved_data["l4"].toInt(), ctx->saved_data["l5"].toInt());
        variable_list out(8);
        out[0] = r[0];
        out[1] = r[1];
        return out;
    }
};

// ------------------------------------------------------------------ force / virial
void expect(const Tensor &t, int64_t numel, const char *op, const char *what) {
    TORCH_CHECK(t.numel() == numel, op, ": ", what, " has ", t.numel(), " elements, expected ", numel);
}

class Force : public torch::autograd::Function<Force> {
  public:
    static Tensor forward(AutogradContext *ctx, Tensor NL, Tensor dE, Tensor Ri_d, Tensor F) {
        const int f64 = f64_of(dE, "calculateNepForce");
        Tensor nl = NL.to(torch::kLong).contiguous();
        Tensor e = dE.contiguous(), d = Ri_d.to(dE.scalar_type()).contiguous();
        Tensor out = F.to(dE.scalar_type()).contiguous().clone();
        const int64_t N = nl.size(0), S = nl.dim() == 2 ? nl.size(1) : 0;
        expect(e, N * S * 4, "calculateNepForce", "dE [N, S, 4]");
        expect(d, N * S * 12, "calculateNepForce", "Ri_d [N, S, 4, 3]");
        expect(out, N * 3, "calculateNepForce", "F [N, 3]");
        if (nl.numel()) TORCH_CHECK(nl.max().item<int64_t>() < N, "calculateNepForce: neighbor index beyond N");
        check(matpl_nep_force(f64, nl.data_ptr<int64_t>(), e.data_ptr(), d.data_ptr(), out.data_ptr(), N, (int)S,
                              stream()),
              "calculateNepForce");
        ctx->save_for_backward({nl, d});
        ctx->saved_data["shape"] = dE.sizes().vec();
        return out;
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        auto sv = ctx->get_saved_variables();
        Tensor nl = sv[0], d = sv[1];
        const int64_t N = nl.size(0), S = nl.size(1);
        Tensor g = torch::empty(ctx->saved_data["shape"].toIntVector(), d.options());
        Tensor gF = gout[0].to(d.scalar_type()).contiguous();
        check(matpl_nep_force_grad(f64_of(d, "calculateNepForce"), nl.data_ptr<int64_t>(), d.data_ptr(),
                                   gF.data_ptr(), g.data_ptr(), N, (int)S, stream()),
              "calculateNepForce backward");
        return {Tensor(), g, Tensor(), Tensor()};
    }
};

// start [F+1] (first atom of each frame) and af [N] (frame of each atom) from num_atom [F]
std::pair<Tensor, Tensor> frame_index(const Tensor &num_atom, const torch::Device &dev) {
    Tensor na = num_atom.reshape(-1).to(torch::kLong).to(dev);
    Tensor start = torch::zeros({na.numel() + 1}, na.options());
    if (na.numel()) start.slice(0, 1).copy_(torch::cumsum(na, 0));
    Tensor af = torch::repeat_interleave(torch::arange(na.numel(), na.options().dtype(torch::kInt)), na);
    return {start, af};
}

class Virial : public torch::autograd::Function<Virial> {
  public:
    static variable_list forward(AutogradContext *ctx, Tensor NL, Tensor dE, Tensor Rij, Tensor Ri_d,
                                 Tensor num_atom) {
        const int f64 = f64_of(dE, "calculateNepVirial");
        Tensor nl = NL.to(torch::kLong).contiguous();
        Tensor e = dE.contiguous(), r = Rij.to(dE.scalar_type()).contiguous();
        Tensor d = Ri_d.to(dE.scalar_type()).contiguous();
        const int64_t N = nl.size(0), S = nl.dim() == 2 ? nl.size(1) : 0;
        expect(e, N * S * 4, "calculateNepVirial", "dE [N, S, 4]");
        expect(d, N * S * 12, "calculateNepVirial", "Ri_d [N, S, 4, 3]");
        expect(r, N * S * 3, "calculateNepVirial", "Rij [N, S, 3]");
        if (nl.numel()) TORCH_CHECK(nl.max().item<int64_t>() < N, "calculateNepVirial: neighbor index beyond N");
        TORCH_CHECK(num_atom.sum().item<int64_t>() == N, "calculateNepVirial: num_atom sums to ",
                    num_atom.sum().item<int64_t>(), ", the list has ", N, " atoms");
        auto fi = frame_index(num_atom, dE.device());
        const int64_t nf = fi.first.numel() - 1;
        Tensor v = torch::empty({nf, 9}, dE.options());
        Tensor av = torch::empty({N, 9}, dE.options());
        check(matpl_nep_virial(f64, nl.data_ptr<int64_t>(), e.data_ptr(), r.data_ptr(), d.data_ptr(),
                               fi.first.data_ptr<int64_t>(), nf, N, (int)S, v.data_ptr(), av.data_ptr(), stream()),
              "calculateNepVirial");
        ctx->save_for_backward({nl, r, d, fi.second});
        ctx->saved_data["shape"] = dE.sizes().vec();
        ctx->mark_non_differentiable({av});
        return {v, av};
    }

    static variable_list backward(AutogradContext *ctx, variable_list gout) {
        if (!gout[0].defined()) return {Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
        auto sv = ctx->get_saved_variables();
        Tensor nl = sv[0], r = sv[1], d = sv[2], af = sv[3];
        const int64_t N = nl.size(0), S = nl.size(1);
        Tensor g = torch::empty(ctx->saved_data["shape"].toIntVector(), d.options());
        Tensor gV = gout[0].to(d.scalar_type()).contiguous();
        check(matpl_nep_virial_grad(f64_of(d, "calculateNepVirial"), nl.data_ptr<int64_t>(), r.data_ptr(),
                                    d.data_ptr(), gV.data_ptr(), af.data_ptr<int32_t>(), g.data_ptr(), N, (int)S,
                                    stream()),
              "calculateNepVirial backward");
        return {Tensor(), g, Tensor(), Tensor(), Tensor()};
    }
};

// ------------------------------------------------------------------ neighbor search
struct Geo {
    Tensor start, af, box, bo, nc, pos;
};

Geo nep_geo(const Tensor &num_atoms, const Tensor &box, const Tensor &box_orig, const Tensor &num_cell,
            const Tensor &position, const char *op) {
    TORCH_CHECK(position.is_cuda(), op, ": tensors must be on the GPU");
    const int64_t nf = num_atoms.numel();
    TORCH_CHECK(position.dim() == 2 && position.size(1) == 3 && num_atoms.sum().item<int64_t>() == position.size(0),
                op, ": position must be [sum(num_atoms), 3]");
    TORCH_CHECK(box.numel() == 18 * nf && box_orig.numel() == 9 * nf && num_cell.numel() == 3 * nf, op,
                ": box [F*18], box_orig [F*9], num_cell [F*3] expected");
    auto fi = frame_index(num_atoms, position.device());
    const auto dev = position.device();
    return {fi.first, fi.second, box.to(dev, torch::kDouble).contiguous(), box_orig.to(dev, torch::kDouble).contiguous(),
            num_cell.to(dev, torch::kLong).contiguous(), position.to(torch::kDouble).contiguous()};
}

}  // namespace

std::vector<Tensor> calculateNepFeat(Tensor coeff2, Tensor d12_radial, Tensor NL_radial, Tensor atom_map,
                                     Tensor feats, double rcut_radial, int64_t feat_multi_nums) {
    (void)feats;
    (void)feat_multi_nums;
    return {Radial::apply(coeff2, d12_radial, NL_radial, atom_map, rcut_radial)};
}

std::vector<Tensor> calculateNepMbFeat(Tensor coeff3, Tensor d12_angluar, Tensor NL_angluar, Tensor atom_map,
                                       Tensor feats, int64_t feat_2b_num, int64_t lmax_3, int64_t lmax_4,
                                       int64_t lmax_5, double rcut_angluar) {
    (void)feats;
    (void)feat_2b_num;
    return {Angular::apply(coeff3, d12_angluar, NL_angluar, atom_map, rcut_angluar, lmax_3, lmax_4, lmax_5)};
}

std::vector<Tensor> calculateNepForce(Tensor list_neigh, Tensor dE, Tensor Ri_d, Tensor F) {
    return {Force::apply(list_neigh, dE, Ri_d, F)};
}

std::vector<Tensor> calculateNepVirial(Tensor list_neigh, Tensor dE, Tensor Rij, Tensor Ri_d, Tensor num_atom) {
    auto r = Virial::apply(list_neigh, dE, Rij, Ri_d, num_atom);
    return {r[0], r[1]};
}

std::vector<Tensor> calculate_maxneigh(const Tensor &num_atoms, const Tensor &box, const Tensor &box_orig,
                                       const Tensor &num_cell, const Tensor &position, double cutoff_2b,
                                       double cutoff_3b, int64_t atom_type_num, const Tensor &atom_type_map,
                                       bool with_type) {
    Geo g = nep_geo(num_atoms, box, box_orig, num_cell, position, "calculate_maxneigh");
    const int64_t N = g.pos.size(0);
    Tensor tm = atom_type_map.to(g.pos.device(), torch::kLong).contiguous();
    if (with_type)
        TORCH_CHECK(tm.numel() == N && (N == 0 || tm.max().item<int64_t>() < atom_type_num),
                    "calculate_maxneigh: atom_type_map must give each atom a type < atom_type_num");
    std::vector<int64_t> shape = with_type ? std::vector<int64_t>{N, atom_type_num} : std::vector<int64_t>{N};
    auto lopt = g.pos.options().dtype(torch::kLong);
    Tensor NN_r = torch::empty(shape, lopt), NN_a = torch::empty(shape, lopt);
    check(matpl_nep_neighbor(g.box.data_ptr<double>(), g.bo.data_ptr<double>(), g.nc.data_ptr<int64_t>(),
                             g.start.data_ptr<int64_t>(), g.af.data_ptr<int32_t>(), tm.data_ptr<int64_t>(),
                             g.pos.data_ptr<double>(), cutoff_2b, cutoff_3b, 0, 0, N, with_type ? 2 : 1,
                             (int)atom_type_num, 1, NN_r.data_ptr<int64_t>(), nullptr, NN_a.data_ptr<int64_t>(),
                             nullptr, nullptr, nullptr, nullptr, stream()),
          "calculate_maxneigh");
    return {NN_r.to(num_atoms.scalar_type()), NN_a.to(num_atoms.scalar_type())};
}

std::vector<Tensor> calculate_neighbor(const Tensor &num_atoms, const Tensor &atom_type_map, const Tensor &atom_types,
                                       const Tensor &box, const Tensor &box_orig, const Tensor &num_cell,
                                       const Tensor &position, double cutoff_2b, double cutoff_3b,
                                       int64_t max_NN_radial, int64_t max_NN_angular, bool with_rij) {
    (void)atom_type_map;
    (void)atom_types;
    Geo g = nep_geo(num_atoms, box, box_orig, num_cell, position, "calculate_neighbor");
    const int64_t N = g.pos.size(0), w = with_rij ? 4 : 3;
    auto lopt = g.pos.options().dtype(torch::kLong);
    Tensor NN_r = torch::empty({N}, lopt), NN_a = torch::empty({N}, lopt);
    Tensor NL_r = torch::empty({N, max_NN_radial}, lopt), NL_a = torch::empty({N, max_NN_angular}, lopt);
    Tensor Ri_r = torch::empty({N, max_NN_radial, w}, g.pos.options());
    Tensor Ri_a = torch::empty({N, max_NN_angular, w}, g.pos.options());
    Tensor status = torch::zeros({1}, g.pos.options().dtype(torch::kInt));
    check(matpl_nep_neighbor(g.box.data_ptr<double>(), g.bo.data_ptr<double>(), g.nc.data_ptr<int64_t>(),
                             g.start.data_ptr<int64_t>(), g.af.data_ptr<int32_t>(), nullptr, g.pos.data_ptr<double>(),
                             cutoff_2b, cutoff_3b, max_NN_radial, max_NN_angular, N, 0, 0, with_rij ? 1 : 0,
                             NN_r.data_ptr<int64_t>(), NL_r.data_ptr<int64_t>(), NN_a.data_ptr<int64_t>(),
                             NL_a.data_ptr<int64_t>(), Ri_r.data_ptr<double>(), Ri_a.data_ptr<double>(),
                             status.data_ptr<int32_t>(), stream()),
          "calculate_neighbor");
    const int s = status.item<int>();
    TORCH_CHECK(s == 0, "calculate_neighbor: more than max_NN neighbors (atom ", s - 1, ")");
    auto it = num_atoms.scalar_type();
    auto dt = position.scalar_type();
    return {NN_r.to(it), NN_a.to(it), NL_r.to(it), NL_a.to(it), Ri_r.to(dt), Ri_a.to(dt)};
}

std::vector<Tensor> calculate_descriptor(const Tensor &weight_radial, const Tensor &weight_angular,
                                         const Tensor &Ri_radial, const Tensor &NL_radial,
                                         const Tensor &atom_type_map, double cutoff_radial, double cutoff_angular,
                                         int64_t max_NN_radial, int64_t lmax_3, int64_t lmax_4, int64_t lmax_5) {
    (void)max_NN_radial;
    Tensor Ri = Ri_radial;
    if (Ri.size(-1) == 3) {
        Tensor r = torch::sqrt((Ri * Ri).sum(-1, true));
        Ri = torch::cat({r, Ri}, -1);
    }
    Ri = Ri.contiguous();
    Tensor NL = NL_radial.to(torch::kLong).contiguous();
    Tensor am = atom_type_map.to(torch::kLong);
    // the angular sums run over the radial list: neighbors beyond cutoff_angular have f_k = 0
    Tensor fr = Radial::apply(weight_radial, Ri, NL, am, cutoff_radial);
    Tensor fa = Angular::apply(weight_angular, Ri, NL, am, cutoff_angular, lmax_3, lmax_4, lmax_5);
    return {torch::cat({fr, fa}, -1)};
}

}  // namespace matpl_torch
