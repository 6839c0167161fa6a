// C++ facade checks, written the way the reference's own example drives its API
// (Q/examples/examples.cu:885-913). Modes:
//   tables <case>  print the facade's term table, group array and characters as
//                  JSON (no GPU); tests compare them with the oracle's restatement
//   solve <case>   run HeisenbergXXZAbelianSymms on the GPU and print per-irrep
//                  energies as JSON
//   expect <case>  ground-state expectation values through ComputeStaticExpValZeroT
//   ref_cuda <case>  the reference free function on reference-layout arrays of
//                  irrep 0 (built through the C ABI): LanczosEigVals_CUDA on device
//                  copies (borrowed device pointers) and LanczosEigVals on the host
//                  arrays, plus a solve through the HeisenbergXXZAbelianSymms_CUDA name
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "qmb/qmb.hpp"

using namespace qmb;
using cd = std::complex<double>;

struct Case {
    Heisenberg<cd> model;
    SymmetryClass<cd> *group;
    std::vector<SymmetryClass<cd> *> owned;
};

static HeisenbergInfo<cd> chain_info(double J1, double delta, double J2) {
    HeisenbergInfo<cd> info;
    for (int d = 1; d <= 2; ++d) {
        const double J = d == 1 ? J1 : J2;
        if (J == 0.0) continue;
        info.Jxy_terms.push_back({d, 0, 0, 0});
        info.Jxy_terms.push_back({-d, 0, 0, 0});
        info.Jxy_couplings.push_back(cd(J / 2.0));
        info.Jxy_couplings.push_back(cd(J / 2.0));
        info.Jz_terms.push_back({d, 0, 0, 0});
        info.Jz_couplings.push_back(J * delta);
    }
    return info;
}

static Case make_case(const std::string &name) {
    Case c;
    if (name == "chain16") {
        c.model = Heisenberg<cd>(create_square_lattice_info(16, 1), chain_info(1.0, 1.0, 0.0));
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        c.owned.push_back(t);
        c.group = t;
    } else if (name == "chain12_mg_flip") {
        c.model = Heisenberg<cd>(create_square_lattice_info(12, 1), chain_info(1.0, 1.0, 0.5));
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        auto *f = new SpinFlipGroup<cd>(*t);
        c.owned = {t, f};
        c.group = f;
    } else if (name == "square2x4_example") {
        // the reference's active example: CreateHeisenbergXXXSquare<complex>(2, 4, {1.0, 0.15})
        c.model = CreateHeisenbergXXXSquare<cd>(2, 4, {1.0, 0.15});
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        c.owned.push_back(t);
        c.group = t;
    } else if (name == "triangular3x3") {
        c.model = CreateHeisenbergXXXTriangular<cd>(3, 3, {1.0});
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        c.owned.push_back(t);
        c.group = t;
    } else if (name == "square4x4_flip") {
        HeisenbergInfo<cd> info;
        info.Jxy_terms = {{1, 0, 0, 0}, {-1, 0, 0, 0}, {0, 1, 0, 0}, {0, -1, 0, 0}};
        info.Jxy_couplings = {cd(0.5), cd(0.5), cd(0.5), cd(0.5)};
        info.Jz_terms = {{1, 0, 0, 0}, {0, 1, 0, 0}};
        info.Jz_couplings = {1.0, 1.0};
        c.model = Heisenberg<cd>(create_square_lattice_info(4, 4), info);
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        auto *f = new SpinFlipGroup<cd>(*t);
        c.owned = {t, f};
        c.group = f;
    } else if (name == "honeycomb2x3_phase") {
        // two sites per cell, complex hopping J/2 exp(+-i 0.37), Jz = 0.8
        LatticeGeometryInfo g;
        const double sq3 = std::sqrt(3.0) / 2.0;
        g.A_mat = {{1.0, 0.5}, {0.0, sq3}};
        g.N1 = 2;
        g.N2 = 3;
        g.rx_alpha = {0.0, 0.5};
        g.ry_alpha = {0.0, sq3 / 3.0};
        HeisenbergInfo<cd> info;
        const cd hop = 0.5 * std::polar(1.0, 0.37);
        info.Jxy_terms = {{0, 0, 1, 0}, {-1, 0, 1, 0}, {0, -1, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 1}, {0, 1, 0, 1}};
        info.Jxy_couplings = {hop, hop, hop, std::conj(hop), std::conj(hop), std::conj(hop)};
        info.Jz_terms = {{0, 0, 1, 0}, {-1, 0, 1, 0}, {0, -1, 1, 0}};
        info.Jz_couplings = {0.8, 0.8, 0.8};
        c.model = Heisenberg<cd>(g, info);
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        c.owned.push_back(t);
        c.group = t;
    } else if (name == "chain12_reflection") {
        // translations x spin flip x the site reflection x -> -x, folded into the ring
        c.model = Heisenberg<cd>(create_square_lattice_info(12, 1), chain_info(1.0, 1.0, 0.4));
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        auto *f = new SpinFlipGroup<cd>(*t);
        auto *p = new SymmetryClassC({{cd(1), cd(1)}, {cd(1), cd(-1)}},
                                     {Iden, [](const std::vector<double> &X, const std::vector<double> &Y) {
                                          return sigma_n_arb(X, Y, 0.0);
                                      }});
        const auto geom = c.model.GetGeomInfo();
        auto *pg = new ProductGroup<cd>(*f, *p, c.model.GetR1(), c.model.GetR2(), &geom);
        c.owned = {t, f, p, pg};
        c.group = pg;
    } else if (name == "square4x4_c4") {
        // translations x spin flip x C4 rotations about site 0, folded into the torus
        HeisenbergInfo<cd> info;
        info.Jxy_terms = {{1, 0, 0, 0}, {-1, 0, 0, 0}, {0, 1, 0, 0}, {0, -1, 0, 0}};
        info.Jxy_couplings = {cd(0.5), cd(0.5), cd(0.5), cd(0.5)};
        info.Jz_terms = {{1, 0, 0, 0}, {0, 1, 0, 0}};
        info.Jz_couplings = {1.0, 1.0};
        c.model = Heisenberg<cd>(create_square_lattice_info(4, 4), info);
        auto *t = new Tnm_group(c.model.GetGeomInfo());
        auto *f = new SpinFlipGroup<cd>(*t);
        auto *r = new Cn_group(4, {0.0, 0.0});
        const auto geom = c.model.GetGeomInfo();
        auto *pg = new ProductGroup<cd>(*f, *r, c.model.GetR1(), c.model.GetR2(), &geom);
        c.owned = {t, f, r, pg};
        c.group = pg;
    } else {
        std::fprintf(stderr, "unknown case %s\n", name.c_str());
        std::exit(2);
    }
    return c;
}

static void print_tables(const Case &c) {
    const auto ov = BuildOperatorVecs(c.model.GetH());
    const int LS = c.model.GetLS();
    const auto el = c.group->GetGroupElemVec();
    std::vector<uint32_t> ge(el.size() * LS);
    c.group->GetGroupElemArr(c.model.GetR1(), c.model.GetR2(), ge.data());
    std::printf("{\"n_sites\": %d, \"H_size\": %d, \"H_col\": %d, \"op\": [", LS, ov.H_size, ov.H_col_size);
    for (size_t i = 0; i < ov.H_decode_table.size(); ++i) std::printf("%s%d", i ? "," : "", ov.H_decode_table[i]);
    std::printf("], \"site\": [");
    for (size_t i = 0; i < ov.H_site_table.size(); ++i) std::printf("%s%d", i ? "," : "", ov.H_site_table[i]);
    std::printf("], \"scalar\": [");
    for (size_t i = 0; i < ov.H_scalar_table.size(); ++i)
        std::printf("%s[%.17g,%.17g]", i ? "," : "", ov.H_scalar_table[i].real(), ov.H_scalar_table[i].imag());
    std::printf("], \"group_el\": [");
    for (size_t i = 0; i < ge.size(); ++i) std::printf("%s%u", i ? "," : "", ge[i]);
    std::printf("], \"flip\": [");
    const auto fl = c.group->GetSpinFlips();
    for (size_t i = 0; i < fl.size(); ++i) std::printf("%s%d", i ? "," : "", fl[i]);
    std::printf("], \"chars\": [");
    const auto ch = c.group->GetCharTable();
    bool first = true;
    for (const auto &row : ch)
        for (const auto &v : row) {
            std::printf("%s[%.17g,%.17g]", first ? "" : ",", v.real(), v.imag());
            first = false;
        }
    std::printf("]}\n");
}

static void solve(const Case &c, bool fixed, bool batched = false, bool locality = false) {
    EDOptions o;
    o.fixed_iter = fixed;
    o.batched = batched;
    o.locality_slots = locality;
    HeisenbergXXZAbelianSymms<cd> ham(c.model, c.group, -1, -1, 1, 2212456, fixed ? 60 : 50, 1e-5, o);
    const auto E = ham.get_E_vecs();
    std::printf("{\"E\": [");
    for (size_t a = 0; a < E.size(); ++a) {
        const double e = E[a][0];
        if (std::isinf(e))
            std::printf("%snull", a ? "," : "");
        else
            std::printf("%s%.15g", a ? "," : "", e);
    }
    std::printf("], \"E_gs\": %.15g, \"GS_irrep\": %d, \"krylov_len\": %zu}\n", ham.GetGroundStateEnergy(),
                ham.GetGroundStateIrrep(), ham.GetKrylovGS().size());
}

// Ground-state expectation values (ComputeStaticExpValZeroT), including the
// reference example's Sz_0 Sz_3 (Q/examples/examples.cu:897-909).
static void expect(const Case &c, bool locality = false) {
    EDOptions o;
    o.locality_slots = locality;
    HeisenbergXXZAbelianSymms<cd> ham(c.model, c.group, -1, -1, 1, 2212456, 50, 1e-5, o);
    Operator<cd> Sz0(0, OperatorType::Sz, 1.0), Sz1(3, OperatorType::Sz, 1.0);
    ManyBodyOperator<cd> szsz;
    szsz.AddTerm({Sz0, Sz1});
    ManyBodyOperator<cd> hop;  // neither Hermitian nor translation invariant
    hop.AddTerm({Operator<cd>(0, OperatorType::Sp, cd(0.4, 0.9)), Operator<cd>(1, OperatorType::Sm, 1.0)});
    const auto geom = c.model.GetGeomInfo();
    const auto Ak = FullFourierOperator(Sz<cd>(0), geom);
    const auto Amk = FullFourierOperator(Sz<cd>(0), geom, true);
    const ManyBodyOperator<cd> szz = Amk[1][0][0] * Ak[1][0][0];  // S^zz at k = b1/N1
    auto put = [](const char *name, cd v, bool last = false) {
        std::printf("\"%s\": [%.17g, %.17g]%s", name, v.real(), v.imag(), last ? "" : ", ");
    };
    std::printf("{\"E_gs\": %.17g, \"GS_irrep\": %d, ", ham.GetGroundStateEnergy(), ham.GetGroundStateIrrep());
    put("szsz", ham.ComputeStaticExpValZeroT(szsz));
    put("hop", ham.ComputeStaticExpValZeroT(hop));
    put("szz", ham.ComputeStaticExpValZeroT(szz));
    put("H", ham.ComputeStaticExpValZeroT(c.model.GetH()), true);
    std::printf("}\n");
}

// The reference's free function on its own row-major arrays (irrep 0, reference
// layout): the driver's row width, device copies for LanczosEigVals_CUDA.
static void ref_cuda(const Case &c) {
    const int LS = c.model.GetLS();
    const auto chars = c.group->GetCharTable();
    const int G = static_cast<int>(c.group->GetGroupElemVec().size()), NIr = static_cast<int>(chars.size());
    std::vector<uint32_t> ge((size_t)G * LS);
    c.group->GetGroupElemArr(c.model.GetR1(), c.model.GetR2(), ge.data());
    auto fl = c.group->GetSpinFlips();
    if (fl.empty()) fl.assign(G, 0);
    std::vector<double> ch((size_t)NIr * G * 2);
    for (int a = 0; a < NIr; ++a)
        for (int g = 0; g < G; ++g) {
            ch[((size_t)a * G + g) * 2] = chars[a][g].real();
            ch[((size_t)a * G + g) * 2 + 1] = chars[a][g].imag();
        }
    const auto ov = BuildOperatorVecs(c.model.GetH());
    std::vector<double> sc((size_t)ov.H_size * 2);
    for (int h = 0; h < ov.H_size; ++h) {
        sc[2 * h] = ov.H_scalar_table[h].real();
        sc[2 * h + 1] = ov.H_scalar_table[h].imag();
    }
    const int n_up = LS / 2, W = ov.H_size * n_up / LS + 1;
    qmb_ctx *ctx = nullptr;
    qmb_group *grp = nullptr;
    qmb_terms *terms = nullptr;
    qmb_basis *b = nullptr;
    qmb_ell *e = nullptr;
    check(qmb_ctx_create(0, &ctx), "qmb_ctx_create");
    check(qmb_group_create(LS, G, ge.data(), fl.data(), NIr, ch.data(), &grp), "qmb_group_create");
    check(qmb_terms_create(LS, ov.H_size, ov.H_col_size, ov.H_decode_table.data(), ov.H_site_table.data(), sc.data(),
                           ov.H_NOTerms.data(), &terms),
          "qmb_terms_create");
    check(qmb_basis_build(ctx, grp, n_up, 0, QMB_LAYOUT_REFERENCE, &b), "qmb_basis_build");
    check(qmb_ell_build(ctx, b, terms, QMB_COMPLEX, W, &e), "qmb_ell_build");
    qmb_ell_info inf;
    check(qmb_ell_get_info(e, &inf), "qmb_ell_get_info");
    const int n = (int)inf.n_rows;
    std::vector<uint32_t> ind((size_t)n * W);
    std::vector<cd> val((size_t)n * W);
    std::vector<int16_t> track(n);
    check(qmb_ell_export_reference(e, ind.data(), reinterpret_cast<double *>(val.data()), track.data()),
          "qmb_ell_export_reference");
    void *d_ind = nullptr, *d_val = nullptr, *d_tr = nullptr;
    if (hipMalloc(&d_ind, ind.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_val, val.size() * sizeof(cd)) != hipSuccess ||
        hipMalloc(&d_tr, track.size() * sizeof(int16_t)) != hipSuccess ||
        hipMemcpy(d_ind, ind.data(), ind.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_val, val.data(), val.size() * sizeof(cd), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tr, track.data(), track.size() * sizeof(int16_t), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("device copies of the reference arrays failed");
    const int NIter = 60;
    double E_dev = 0.0, E_host = 0.0;
    std::vector<double> kr_dev(NIter), kr_host(NIter);
    const double r_dev = LanczosEigVals_CUDA<cd>(&E_dev, 1, NIter, n, W, 2212456, static_cast<const uint32_t *>(d_ind),
                                                 static_cast<const cd *>(d_val), static_cast<const int16_t *>(d_tr),
                                                 kr_dev.data(), true, true);
    const double r_host = LanczosEigVals<cd>(&E_host, 1, NIter, n, W, 2212456, ind.data(), val.data(), track.data(),
                                             kr_host.data(), true, true);
    (void)hipFree(d_ind);
    (void)hipFree(d_val);
    (void)hipFree(d_tr);
    qmb_ell_destroy(e);
    qmb_basis_destroy(b);
    qmb_terms_destroy(terms);
    qmb_group_destroy(grp);
    qmb_ctx_destroy(ctx);
    // the driver under the reference's class name
    HeisenbergXXZAbelianSymms_CUDA<cd> ham(c.model, c.group);
    bool same_kr = true;
    for (int i = 0; i < NIter; ++i) same_kr = same_kr && kr_dev[i] == kr_host[i];
    std::printf("{\"n\": %d, \"W\": %d, \"E_dev\": %.17g, \"E_host\": %.17g, \"res_dev\": %.17g, "
                "\"res_host\": %.17g, \"same_krylov\": %s, \"E_driver_irrep0\": %.17g}\n",
                n, W, E_dev, E_host, r_dev, r_host, same_kr ? "true" : "false", ham.get_E_vecs()[0][0]);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr,
                     "usage: %s tables|solve|solve_fixed|solve_batched|solve_locality|expect|expect_locality|"
                     "ref_cuda <case>\n",
                     argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    Case c = make_case(argv[2]);
    try {
        if (mode == "tables")
            print_tables(c);
        else if (mode == "solve")
            solve(c, false);
        else if (mode == "solve_fixed")
            solve(c, true);
        else if (mode == "solve_batched")
            solve(c, false, true);
        else if (mode == "solve_locality")
            solve(c, false, false, true);
        else if (mode == "expect_locality")
            expect(c, true);
        else if (mode == "expect")
            expect(c);
        else if (mode == "ref_cuda")
            ref_cuda(c);
        else
            return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    for (auto *g : c.owned) delete g;
    return 0;
}
