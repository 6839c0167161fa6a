// Quantum operators of the C++ facade: the reference's OperatorType, Operator<T>,
// ManyBodyOperator<T> (sum of products) and the Sz/Sp/Sm/SzSz helpers.
// Reference: Q/quantum_operators/SingleParticleOperators.h:4-29,
// SingleParticleOperators.cu:7-90, ArbitraryOperator.h:7-80,
// ArbitraryOperator.cu:9-242 (Q/ = QMBCuda/). T is double or
// std::complex<double> (the reference uses float / thrust::complex<float>).
#pragma once

#include <complex>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

namespace qmb {

// Same enumerators in the same order (SingleParticleOperators.h:4); the C ABI
// accepts Sz, Sp and Sm.
enum class OperatorType { Sz, Sp, Sm, c_up, c_dag_up, c_down, c_dag_down, b, b_dag, I };

inline OperatorType GetConjugateOperatorType(OperatorType op) {
    switch (op) {
        case OperatorType::Sp: return OperatorType::Sm;
        case OperatorType::Sm: return OperatorType::Sp;
        case OperatorType::b: return OperatorType::b_dag;
        case OperatorType::b_dag: return OperatorType::b;
        case OperatorType::c_up: return OperatorType::c_dag_up;
        case OperatorType::c_dag_up: return OperatorType::c_up;
        case OperatorType::c_down: return OperatorType::c_dag_down;
        case OperatorType::c_dag_down: return OperatorType::c_down;
        default: return op;  // Sz, I
    }
}

inline std::string PrintOperatorType(OperatorType op) {
    static const char *names[] = {"Sz", "Sp", "Sm", "c_up", "c_dag_up", "c_dw", "c_dag_dw", "b", "b_dag", "I"};
    return names[static_cast<int>(op)];
}

template <class T>
inline T conj_value(const T &x) {
    if constexpr (std::is_same_v<T, std::complex<double>>)
        return std::conj(x);
    else
        return x;
}

template <class T>
class Operator {
  public:
    Operator(int site_in, OperatorType type_in, T scalar_in = T(1.0)) : site(site_in), scalar(scalar_in), type(type_in) {}
    int GetSite() const { return site; }
    T GetScalar() const { return scalar; }
    OperatorType GetType() const { return type; }
    void SetScalar(T s) { scalar = s; }
    void SetSite(int s) { site = s; }

  protected:
    int site;
    T scalar;
    OperatorType type;
};

// Sum of products of single-site operators; a product's scalar is the product
// of its operators' scalars (ArbitraryOperator.cu:124-201).
template <class T>
class ManyBodyOperator {
  public:
    using Term = std::vector<Operator<T>>;
    ManyBodyOperator() = default;
    explicit ManyBodyOperator(std::vector<Term> e) : elements(std::move(e)) {}
    virtual ~ManyBodyOperator() = default;

    std::vector<Term> GetElems() const { return elements; }
    int GetMaxTerms() const {
        size_t m = 0;
        for (const auto &t : elements) m = t.size() > m ? t.size() : m;
        return static_cast<int>(m);
    }
    void AddTerm(Term t) { elements.emplace_back(std::move(t)); }
    void AddSingleOperator(Operator<T> op) { elements.push_back(Term{op}); }
    void LocalizeOperator(int new_site) {
        for (auto &t : elements)
            for (auto &o : t) o.SetSite(new_site);
    }
    void PrintOperator() const {
        for (size_t i = 0; i < elements.size(); ++i) {
            T s = T(1.0);
            for (const auto &o : elements[i]) s *= o.GetScalar();
            if constexpr (std::is_same_v<T, double>)
                std::printf("%f ", s);
            else
                std::printf("(%f%+fi) ", s.real(), s.imag());
            for (const auto &o : elements[i]) std::printf("%s(r=%d)", PrintOperatorType(o.GetType()).c_str(), o.GetSite());
            if (i + 1 < elements.size()) std::printf(" + ");
        }
        std::printf("\n");
    }

    // Concatenation of the term lists (ArbitraryOperator.cu:124-131).
    ManyBodyOperator operator+(const ManyBodyOperator &A) const {
        auto e = elements;
        for (const auto &t : A.elements) e.push_back(t);
        return ManyBodyOperator(e);
    }
    // Distributive product: every term of *this followed by every term of A (:136-149).
    ManyBodyOperator operator*(const ManyBodyOperator &A) const {
        std::vector<Term> e;
        for (const auto &a : elements)
            for (const auto &b : A.elements) {
                Term t = a;
                t.insert(t.end(), b.begin(), b.end());
                e.push_back(t);
            }
        return ManyBodyOperator(e);
    }
    // Scalar multiplication scales the first operator of every term (:153-201).
    ManyBodyOperator operator*(T scalar) const {
        auto e = elements;
        for (auto &t : e)
            if (!t.empty()) t[0].SetScalar(t[0].GetScalar() * scalar);
        return ManyBodyOperator(e);
    }

  protected:
    std::vector<Term> elements;
};

template <class T>
ManyBodyOperator<T> operator*(T scalar, const ManyBodyOperator<T> &A) {
    return A * scalar;
}

template <class T>
class Sz : public ManyBodyOperator<T> {
  public:
    explicit Sz(int i) { this->elements.push_back({Operator<T>(i, OperatorType::Sz, T(1.0))}); }
};

template <class T>
class Sp : public ManyBodyOperator<T> {
  public:
    explicit Sp(int i) { this->elements.push_back({Operator<T>(i, OperatorType::Sp, T(1.0))}); }
};

template <class T>
class Sm : public ManyBodyOperator<T> {
  public:
    explicit Sm(int i) { this->elements.push_back({Operator<T>(i, OperatorType::Sm, T(1.0))}); }
};

template <class T>
class SzSz_correlator : public ManyBodyOperator<T> {
  public:
    SzSz_correlator(int i, int j) {
        this->elements.push_back({Operator<T>(i, OperatorType::Sz, T(1.0)), Operator<T>(j, OperatorType::Sz, T(1.0))});
    }
};

// Flat conjugated tables consumed by the device (BuildOperatorVecs,
// Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336):
// op types conjugated, scalar = product of conjugated operator scalars.
template <class T>
struct OperatorVectors {
    std::vector<int32_t> H_decode_table;  // OperatorType as int
    std::vector<std::complex<double>> H_scalar_table;
    std::vector<int32_t> H_site_table;    // -1 pads
    std::vector<int32_t> H_NOTerms;
    int H_size = 0;
    int H_col_size = 0;
};

template <class T>
OperatorVectors<T> BuildOperatorVecs(const ManyBodyOperator<T> &H) {
    OperatorVectors<T> v;
    const auto el = H.GetElems();
    v.H_size = static_cast<int>(el.size());
    v.H_col_size = H.GetMaxTerms();
    v.H_decode_table.assign((size_t)v.H_size * v.H_col_size, 0);
    v.H_site_table.assign((size_t)v.H_size * v.H_col_size, -1);
    v.H_scalar_table.assign(v.H_size, std::complex<double>(1.0, 0.0));
    v.H_NOTerms.assign(v.H_size, 0);
    for (int h = 0; h < v.H_size; ++h) {
        v.H_NOTerms[h] = static_cast<int>(el[h].size());
        for (size_t k = 0; k < el[h].size(); ++k) {
            v.H_decode_table[(size_t)h * v.H_col_size + k] = static_cast<int>(GetConjugateOperatorType(el[h][k].GetType()));
            v.H_site_table[(size_t)h * v.H_col_size + k] = el[h][k].GetSite();
            v.H_scalar_table[h] *= std::complex<double>(conj_value(el[h][k].GetScalar()));
        }
    }
    return v;
}

}  // namespace qmb
