// ED driver and Lanczos free function of the C++ facade, on top of the C ABI.
// Reference: Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.h:27-89
// and .cu:20-251 (driver), Q/exact_diagonalization/lanczos.h:21-32 (free function).
//
// HeisenbergXXZAbelianSymms<T> keeps the reference constructor and getters.
// Differences, all documented at the ABI: FP64 results; states may have up to
// 64 sites; Lanczos runs to a tolerance (EDOptions::fixed_iter restores the
// reference's fixed NIter + residual check + reseed loop); groups may carry
// spin flips; the compact layout drops null orbits (the reference keeps them as
// empty rows; EDOptions::reference_layout restores that).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../qmb.h"
#include "lattice.hpp"
#include "operators.hpp"
#include "symmetry.hpp"

namespace qmb {

struct Status : std::runtime_error {
    qmb_status code;
    Status(qmb_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(qmb_status s, const char *what) {
    if (s != QMB_OK) throw Status(s, std::string(what) + ": " + qmb_last_error());
}

struct EDOptions {
    int irrep = -1;                // -1: every irrep (the reference's loop); else only this one
    bool reference_layout = false; // keep null orbits as empty rows
    bool fixed_iter = false;       // exactly NIter steps, residual check and reseeding (reference semantics)
    double tol = 1e-13;            // relative eigenvalue tolerance when not fixed_iter
    int max_iter = 2000;           // iteration cap when not fixed_iter
    int device = 0;
    bool batched = false;          // not fixed_iter: solve all irreps in one qmb_lanczos_batched call
                                   // (the ground-state irrep is then re-solved alone for its Krylov vector)
    bool locality_slots = false;   // QMB_LOCALITY_SLOTS: row slots in locality order (same matrix; no
                                   // counterpart in the reference, whose slots follow the term table)
};

template <class T>
class HeisenbergXXZAbelianSymms {
  public:
    HeisenbergXXZAbelianSymms(Heisenberg<T> model, SymmetryClass<T> *group, int hop_f = -1, int GS_sector = -1,
                              int NStates = 1, long long seed = 2212456, int NIter = 50, double tol_norm = 1e-5,
                              EDOptions opts = EDOptions())
        : lattice_model(std::move(model)), SymmGroup(group), hop_f_(hop_f), NStates_(NStates), NIter_(NIter),
          tol_norm_(tol_norm), opts_(opts) {
        if (NStates != 1) throw Status(QMB_ERR_INVALID, "NStates must be 1 (as in the reference)");
        LS = lattice_model.GetLS();
        GS_sector_ = GS_sector < 0 ? LS / 2 : GS_sector;
        const auto chars = group->GetCharTable();
        const auto elems = group->GetGroupElemVec();
        auto flips = group->GetSpinFlips();
        const int G = static_cast<int>(elems.size()), NIr = static_cast<int>(chars.size());
        group_el_arr.assign((size_t)G * LS, 0);
        group->GetGroupElemArr(lattice_model.GetR1(), lattice_model.GetR2(), group_el_arr.data());
        std::vector<double> ch((size_t)NIr * G * 2);
        for (int a = 0; a < NIr; ++a)
            for (int g = 0; g < G; ++g) {
                const std::complex<double> c(chars[a][g]);
                ch[((size_t)a * G + g) * 2] = c.real();
                ch[((size_t)a * G + g) * 2 + 1] = c.imag();
            }
        if (flips.empty()) flips.assign(G, 0);
        ch_ = ch;
        flips_ = flips;
        const OperatorVectors<T> ov = BuildOperatorVecs(lattice_model.GetH());
        std::vector<double> sc((size_t)ov.H_size * 2);
        for (int h = 0; h < ov.H_size; ++h) {
            sc[2 * h] = ov.H_scalar_table[h].real();
            sc[2 * h + 1] = ov.H_scalar_table[h].imag();
        }
        // the reference's row width (…cu:71-72)
        max_terms = hop_f > 0 ? GS_sector_ * hop_f : ov.H_size * GS_sector_ / LS + 1;

        qmb_ctx *ctx = nullptr;
        qmb_group *grp = nullptr;
        qmb_terms *terms = nullptr;
        check(qmb_ctx_create(opts_.device, &ctx), "qmb_ctx_create");
        try {
            check(qmb_group_create(LS, G, group_el_arr.data(), flips.data(), NIr, ch.data(), &grp), "qmb_group_create");
            check(qmb_terms_create(LS, ov.H_size, ov.H_col_size, ov.H_decode_table.data(), ov.H_site_table.data(),
                                   sc.data(), ov.H_NOTerms.data(), &terms),
                  "qmb_terms_create");
            const int a0 = opts_.irrep < 0 ? 0 : opts_.irrep;
            const int a1 = opts_.irrep < 0 ? NIr : opts_.irrep + 1;
            if (opts_.batched && !opts_.fixed_iter) {
                solve_batched(ctx, grp, terms, a0, a1, NIr, ch, G, seed);
            } else {
                for (int a = 0; a < NIr; ++a) {
                    if (a < a0 || a >= a1) {
                        E_vecs.push_back({std::numeric_limits<double>::infinity()});
                        seed_vecs.push_back({seed});
                        KrylovGS_vecs.push_back({});
                        continue;
                    }
                    solve_irrep(ctx, grp, terms, a, ch, G, seed);
                }
            }
        } catch (...) {
            qmb_terms_destroy(terms);
            qmb_group_destroy(grp);
            qmb_ctx_destroy(ctx);
            throw;
        }
        qmb_terms_destroy(terms);
        qmb_group_destroy(grp);
        qmb_ctx_destroy(ctx);
        E_gs = E_vecs[0][0];
        GS_irrep = 0;
        GS_seed = seed_vecs[0][0];
        for (int a = 1; a < NIr; ++a)
            if (E_vecs[a][0] < E_gs) {
                E_gs = E_vecs[a][0];
                GS_irrep = a;
                GS_seed = seed_vecs[a][0];
            }
        KrylovGS_vec = KrylovGS_vecs[GS_irrep];
    }
    virtual ~HeisenbergXXZAbelianSymms() = default;

    std::vector<std::vector<double>> get_E_vecs() const { return E_vecs; }
    std::vector<std::vector<long long>> get_seed_vecs() const { return seed_vecs; }
    double GetGroundStateEnergy() const { return E_gs; }
    int GetGroundStateIrrep() const { return GS_irrep; }
    long long GetGroundStateSeed() const { return GS_seed; }
    std::vector<double> GetKrylovGS() const { return KrylovGS_vec; }
    std::vector<int64_t> get_sector_dims() const { return sector_dims; }
    std::vector<int> get_iterations() const { return iterations; }

    // <Psi|A|Psi> / <Psi|Psi> in the ground state (ComputeStaticExpValZeroT,
    // Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:421-601). Like the
    // reference, it rebuilds the ground-state irrep's basis and Hamiltonian and regenerates
    // the ground state from its seed, then evaluates qmb_expectation on the device. A is any
    // Sz-conserving operator; it need not commute with the group. max_terms <= 0: the
    // Hamiltonian's row width. The reference returns its sum without normalising by
    // <Psi|Psi> (its regenerated vector has norm close to 1); this returns the normalised value.
    std::complex<double> ComputeStaticExpValZeroT(const ManyBodyOperator<std::complex<double>> &A,
                                                  int max_terms_in = 0) const {
        struct Handles {
            qmb_ctx *ctx = nullptr;
            qmb_group *grp = nullptr;
            qmb_terms *th = nullptr, *ta = nullptr;
            qmb_basis *b = nullptr;
            qmb_ell *e = nullptr;
            ~Handles() {
                qmb_ell_destroy(e);
                qmb_basis_destroy(b);
                qmb_terms_destroy(ta);
                qmb_terms_destroy(th);
                qmb_group_destroy(grp);
                qmb_ctx_destroy(ctx);
            }
        } h;
        const int G = static_cast<int>(flips_.size());
        const int NIr = static_cast<int>(ch_.size() / (2 * (size_t)G));
        check(qmb_ctx_create(opts_.device, &h.ctx), "qmb_ctx_create");
        check(qmb_group_create(LS, G, group_el_arr.data(), flips_.data(), NIr, ch_.data(), &h.grp),
              "qmb_group_create");
        auto make_terms = [&](const auto &ov, qmb_terms **out) {
            std::vector<double> sc((size_t)ov.H_size * 2);
            for (int i = 0; i < ov.H_size; ++i) {
                sc[2 * i] = std::complex<double>(ov.H_scalar_table[i]).real();
                sc[2 * i + 1] = std::complex<double>(ov.H_scalar_table[i]).imag();
            }
            check(qmb_terms_create(LS, ov.H_size, ov.H_col_size, ov.H_decode_table.data(), ov.H_site_table.data(),
                                   sc.data(), ov.H_NOTerms.data(), out),
                  "qmb_terms_create");
        };
        make_terms(BuildOperatorVecs(lattice_model.GetH()), &h.th);
        make_terms(BuildOperatorVecs(A), &h.ta);
        const bool real = irrep_is_real(GS_irrep, ch_, G);
        check(qmb_basis_build(h.ctx, h.grp, GS_sector_, GS_irrep,
                              opts_.reference_layout ? QMB_LAYOUT_REFERENCE : QMB_LAYOUT_COMPACT, &h.b),
              "qmb_basis_build");
        qmb_basis_info bi;
        qmb_basis_get_info(h.b, &bi);
        if (bi.n_rows == 0) throw Status(QMB_ERR_INVALID, "ComputeStaticExpValZeroT: no ground state");
        check(qmb_ell_build(h.ctx, h.b, h.th, (real ? QMB_REAL : QMB_COMPLEX) | slot_flags(),
                            max_terms_in > 0 ? max_terms_in : max_terms, &h.e),
              "qmb_ell_build");
        qmb_lanczos_opts o;
        qmb_lanczos_default_opts(&o);
        o.seed = (uint64_t)GS_seed;
        o.want_eigvec = 1;
        if (opts_.fixed_iter) {
            o.max_iter = NIter_;
            o.fixed_iter = 1;
        } else {
            o.max_iter = opts_.max_iter;
            o.tol = opts_.tol;
        }
        std::vector<double> x((size_t)bi.n_rows * (real ? 1 : 2));
        double E = 0.0;
        qmb_lanczos_report rep;
        const qmb_status st = qmb_lanczos(h.ctx, h.e, &o, &E, nullptr, x.data(), &rep);
        if (st != QMB_OK && !(opts_.fixed_iter && st == QMB_ERR_NOT_CONVERGED)) check(st, "qmb_lanczos");
        double out[2] = {0.0, 0.0};
        check(qmb_expectation(h.ctx, h.b, h.ta, real ? 0 : 1, x.data(), 0, 0, out, nullptr), "qmb_expectation");
        double nn = 0.0;
        for (double v : x) nn += v * v;
        return std::complex<double>(out[0], out[1]) / nn;
    }

  protected:
    bool irrep_is_real(int a, const std::vector<double> &ch, int G) const {
        bool real = true;
        for (int g = 0; g < G; ++g) real = real && ch[((size_t)a * G + g) * 2 + 1] == 0.0;
        if constexpr (std::is_same_v<T, std::complex<double>>) {
            for (const auto &t : lattice_model.GetH().GetElems()) {
                std::complex<double> s(1.0);
                for (const auto &o : t) s *= o.GetScalar();
                real = real && s.imag() == 0.0;
            }
        }
        return real;
    }

    // Every irrep in [a0, a1) in one batched solve; the ground-state irrep is then
    // solved alone once more for its Krylov coefficients (batched solves keep none).
    void solve_batched(qmb_ctx *ctx, qmb_group *grp, qmb_terms *terms, int a0, int a1, int NIr,
                       const std::vector<double> &ch, int G, long long seed) {
        std::vector<qmb_basis *> bases(NIr, nullptr);
        std::vector<qmb_ell *> ells(NIr, nullptr);
        auto cleanup = [&]() {
            for (auto *e : ells) qmb_ell_destroy(e);
            for (auto *b : bases) qmb_basis_destroy(b);
        };
        try {
            std::vector<const qmb_ell *> batch;
            std::vector<int> which;
            E_vecs.assign(NIr, {std::numeric_limits<double>::infinity()});
            seed_vecs.assign(NIr, {seed});
            KrylovGS_vecs.assign(NIr, {});
            iterations.assign(NIr, 0);
            sector_dims.clear();
            for (int a = a0; a < a1; ++a) {
                check(qmb_basis_build(ctx, grp, GS_sector_, a,
                                      opts_.reference_layout ? QMB_LAYOUT_REFERENCE : QMB_LAYOUT_COMPACT, &bases[a]),
                      "qmb_basis_build");
                qmb_basis_info bi;
                qmb_basis_get_info(bases[a], &bi);
                sector_dims.push_back(bi.n_rows - bi.n_null);
                if (bi.n_rows == 0) continue;
                check(qmb_ell_build(ctx, bases[a], terms,
                                    (irrep_is_real(a, ch, G) ? QMB_REAL : QMB_COMPLEX) | slot_flags(), max_terms,
                                    &ells[a]),
                      "qmb_ell_build");
                batch.push_back(ells[a]);
                which.push_back(a);
            }
            if (!batch.empty()) {
                qmb_lanczos_opts o;
                qmb_lanczos_default_opts(&o);
                o.seed = (uint64_t)seed;
                o.max_iter = opts_.max_iter;
                o.tol = opts_.tol;
                o.check_every = 5;  // iterations per graph replay
                std::vector<double> E(batch.size());
                std::vector<qmb_lanczos_report> reps(batch.size());
                check(qmb_lanczos_batched(ctx, (int)batch.size(), batch.data(), &o, E.data(), reps.data()),
                      "qmb_lanczos_batched");
                int best = 0;
                for (size_t i = 0; i < batch.size(); ++i) {
                    E_vecs[which[i]] = {E[i]};
                    iterations[which[i]] = reps[i].iters;
                    if (E[i] < E[best]) best = (int)i;
                }
                // Krylov coefficients of the ground-state irrep
                std::vector<double> kr(o.max_iter, 0.0);
                qmb_lanczos_report rep;
                double Eg = 0.0;
                check(qmb_lanczos(ctx, batch[best], &o, &Eg, kr.data(), nullptr, &rep), "qmb_lanczos");
                kr.resize(rep.iters);
                KrylovGS_vecs[which[best]] = kr;
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    }

    void solve_irrep(qmb_ctx *ctx, qmb_group *grp, qmb_terms *terms, int a, const std::vector<double> &ch, int G,
                     long long seed) {
        const bool real = irrep_is_real(a, ch, G);
        qmb_basis *b = nullptr;
        qmb_ell *e = nullptr;
        check(qmb_basis_build(ctx, grp, GS_sector_, a,
                              opts_.reference_layout ? QMB_LAYOUT_REFERENCE : QMB_LAYOUT_COMPACT, &b),
              "qmb_basis_build");
        qmb_basis_info bi;
        qmb_basis_get_info(b, &bi);
        sector_dims.push_back(bi.n_rows - bi.n_null);
        if (bi.n_rows == 0) {  // no orbit survives: the irrep does not occur in this sector
            qmb_basis_destroy(b);
            E_vecs.push_back({std::numeric_limits<double>::infinity()});
            seed_vecs.push_back({seed});
            KrylovGS_vecs.push_back({});
            iterations.push_back(0);
            return;
        }
        qmb_status st = qmb_ell_build(ctx, b, terms, (real ? QMB_REAL : QMB_COMPLEX) | slot_flags(), max_terms, &e);
        if (st != QMB_OK) {
            qmb_basis_destroy(b);
            check(st, "qmb_ell_build");
        }
        qmb_lanczos_opts o;
        qmb_lanczos_default_opts(&o);
        o.seed = (uint64_t)seed;
        std::vector<double> kr;
        double E = 0.0;
        long long used_seed = seed;
        if (opts_.fixed_iter) {
            // the reference's loop: fixed NIter, residual/nobv > 1e-4 -> reseed, at most 10 tries
            // (…cu:189-213); reseeding uses a fixed LCG instead of rand()
            o.max_iter = NIter_;
            o.fixed_iter = 1;
            o.want_eigvec = 1;
            kr.assign(NIter_, 0.0);
            uint64_t lcg = (uint64_t)seed;
            for (int attempt = 0; attempt < 10; ++attempt) {
                qmb_lanczos_report rep;
                o.seed = (uint64_t)used_seed;
                const qmb_status s = qmb_lanczos(ctx, e, &o, &E, kr.data(), nullptr, &rep);
                if (s != QMB_OK && s != QMB_ERR_NOT_CONVERGED) {
                    qmb_ell_destroy(e);
                    qmb_basis_destroy(b);
                    check(s, "qmb_lanczos");
                }
                if (rep.eig_residual / (double)bi.n_rows <= 1e-4) break;
                if (attempt == 9) E = 1e7;  // the reference's "not found" sentinel
                lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
                used_seed = (long long)(lcg >> 33);
            }
            iterations.push_back(NIter_);
        } else {
            o.max_iter = opts_.max_iter;
            o.tol = opts_.tol;
            kr.assign(o.max_iter, 0.0);
            qmb_lanczos_report rep;
            const qmb_status s = qmb_lanczos(ctx, e, &o, &E, kr.data(), nullptr, &rep);
            if (s != QMB_OK) {
                qmb_ell_destroy(e);
                qmb_basis_destroy(b);
                check(s, "qmb_lanczos");
            }
            kr.resize(rep.iters);
            iterations.push_back(rep.iters);
        }
        qmb_ell_destroy(e);
        qmb_basis_destroy(b);
        E_vecs.push_back({E});
        seed_vecs.push_back({used_seed});
        KrylovGS_vecs.push_back(kr);
    }

    Heisenberg<T> lattice_model;
    int LS = 0;
    int GS_sector_ = 0;
    SymmetryClass<T> *SymmGroup;
    int hop_f_, NStates_, NIter_;
    double tol_norm_;
    EDOptions opts_;
    int slot_flags() const { return opts_.locality_slots ? QMB_LOCALITY_SLOTS : 0; }
    std::vector<double> ch_;      // characters [NIr][G] (re, im)
    std::vector<uint8_t> flips_;  // spin flip of each element
    int max_terms = 0;
    std::vector<uint32_t> group_el_arr;
    std::vector<std::vector<double>> E_vecs;
    std::vector<std::vector<long long>> seed_vecs;
    std::vector<std::vector<double>> KrylovGS_vecs;
    std::vector<double> KrylovGS_vec;
    std::vector<int64_t> sector_dims;
    std::vector<int> iterations;
    int GS_irrep = 0;
    double E_gs = 0.0;
    long long GS_seed = 0;
};

// The reference's free function (lanczos.h:21-32) on host reference-layout
// arrays: fixed NIter steps; E_tmp = lowest eigenvalue; returns ||Hx-Ex||/nobv
// when checkEigState. T = double or std::complex<double>.
template <class T>
double LanczosEigVals(double *E_tmp, int NStates, int NIter, int nobv, int max_terms, long long seed,
                      const uint32_t *index_mat, const T *vals_mat, const int16_t *track_vec,
                      double *eig_vec_Krylov_out, bool exportKrylovVec, bool checkEigState) {
    qmb_status st = QMB_OK;
    const double r = qmb_lanczos_eigvals_ref(E_tmp, NStates, NIter, nobv, max_terms, seed, index_mat,
                                             reinterpret_cast<const double *>(vals_mat), track_vec,
                                             eig_vec_Krylov_out, exportKrylovVec ? 1 : 0, checkEigState ? 1 : 0,
                                             std::is_same_v<T, std::complex<double>> ? 1 : 0, &st);
    check(st, "LanczosEigVals");
    return r;
}

// LanczosEigVals_CUDA (lanczos.h:21-32) with the reference's ownership: the ELL
// arrays are borrowed DEVICE pointers (of device 0), converted on the device; E_tmp
// and eig_vec_Krylov_out are host pointers. T = double or std::complex<double>.
template <class T>
double LanczosEigVals_CUDA(double *E_tmp, int NStates, const int NIter, int nobv, int max_terms, long long seed,
                           const uint32_t *index_mat, const T *vals_mat, const int16_t *track_vec,
                           double *eig_vec_Krylov_out, bool exportKrylovVec, bool checkEigState) {
    qmb_status st = QMB_OK;
    const double r = qmb_lanczos_eigvals_ref_device(E_tmp, NStates, NIter, nobv, max_terms, seed, index_mat,
                                                    reinterpret_cast<const double *>(vals_mat), track_vec,
                                                    eig_vec_Krylov_out, exportKrylovVec ? 1 : 0,
                                                    checkEigState ? 1 : 0,
                                                    std::is_same_v<T, std::complex<double>> ? 1 : 0, &st);
    check(st, "LanczosEigVals_CUDA");
    return r;
}

// The reference's driver class name (HeisenbergGenericXXZ_AbelianSymms_CUDA.h:27-50).
template <class T>
using HeisenbergXXZAbelianSymms_CUDA = HeisenbergXXZAbelianSymms<T>;

}  // namespace qmb
