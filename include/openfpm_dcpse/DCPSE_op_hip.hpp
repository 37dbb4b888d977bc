// Operator classes of the reference's DCPSE_op.hpp over the HIP Dcpse_type
// plug point, and the field expressions they are used in:
//
//   auto verletList = domain.getVerlet<VL_NON_SYMMETRIC | VL_SKIP_REF_PART>(rCut);
//   Derivative_x<decltype(verletList)> Dx(domain, verletList, 2, rCut);
//   Derivative_y<decltype(verletList)> Dy(domain, verletList, 2, rCut);
//   auto v = getV<1>(domain);
//   auto P = getV<0>(domain);
//   v = 2 * Dx(P) + Dy(P);
//
// Reference: Derivative_x_T (src/DCPSE/DCPSE_op/DCPSE_op.hpp:649-786) and its
// siblings, Laplacian_T (:1350-1490), Derivative_G_T (:3386-3470) and the
// aliases (:3485-3500). The reference evaluates an assignment particle by
// particle on the host (vector_dist_expression_op<..., VECT_DCPSE>::value,
// DCPSE_op.hpp:65-67); here an assignment is a sum of terms c_k D_k f_k +
// d_j g_j evaluated as whole-field device calls: one dcpse_apply per operator
// term (alpha = c_k, beta accumulating) and one dcpse_axpby per plain field.
// The Laplacian is one operator with the fused signature sum_d D^{2 e_d}
// (one kernel set instead of dim of them; the same values, see Dcpse_hip).
#pragma once

#include <memory>
#include <type_traits>
#include <vector>

#include "Dcpse_hip.hpp"

namespace openfpm_dcpse {

template <unsigned int prp, typename V>
struct field_ref;

// Sum of operator terms c * D(f) and field terms d * g over one particle set.
template <typename V>
struct field_expr {
    typedef typename V::prop_type P;
    struct op_term {
        field_operator<P>* op;
        unsigned int prp;
        double c;
    };
    struct field_term {
        unsigned int prp;
        double c;
    };
    V* parts = nullptr;
    std::vector<op_term> ops;
    std::vector<field_term> fields;

    field_expr scaled(double s) const {
        field_expr e = *this;
        for (auto& t : e.ops) t.c *= s;
        for (auto& t : e.fields) t.c *= s;
        return e;
    }
    field_expr operator-() const { return scaled(-1.0); }
    bool reads(unsigned int prp) const {
        for (auto& t : ops)
            if (t.prp == prp) return true;
        for (auto& t : fields)
            if (t.prp == prp) return true;
        return false;
    }

    // Evaluates the expression into property `out` on the device.
    void evaluate_into(unsigned int out) const {
        V& v = *parts;
        Context& c = context();
        const size_t n = v.size_local();
        P* dst_final = v.devProp(out);
        // an input that is also the output is read from an unmodified copy
        device_array<P> tmp;
        P* dst = dst_final;
        if (reads(out)) {
            tmp.resize(n);
            dst = tmp.data();
        }
        bool first = true;
        for (auto& t : ops) {
            t.op->apply(v.devProp(t.prp), dst, (P)t.c, first ? P(0) : P(1));
            first = false;
        }
        for (auto& t : fields) {
            axpby(c, (int64_t)n, (P)t.c, v.devProp(t.prp), first ? P(0) : P(1), dst);
            first = false;
        }
        if (first && n) hip_check(hipMemset(dst, 0, n * sizeof(P)), "hipMemset");
        if (dst != dst_final && n)
            hip_check(hipMemcpy(dst_final, dst, n * sizeof(P), hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        v.devPropWritten(out);
    }

   private:
    static void axpby(Context& c, int64_t n, float a, const float* x, float b, float* y) {
        c.check(dcpse_axpby_f32(c.get(), n, a, x, b, y), "dcpse_axpby_f32");
    }
    static void axpby(Context& c, int64_t n, double a, const double* x, double b, double* y) {
        c.check(dcpse_axpby_f64(c.get(), n, a, x, b, y), "dcpse_axpby_f64");
    }
};

template <unsigned int prp, typename V>
struct field_ref {
    typedef V vtype;
    V& parts;
    explicit field_ref(V& p) : parts(p) {}

    field_expr<V> as_expr() const {
        field_expr<V> e;
        e.parts = &parts;
        e.fields.push_back({prp, 1.0});
        return e;
    }
    operator field_expr<V>() const { return as_expr(); }

    field_ref& operator=(const field_expr<V>& e) {
        if (e.parts != nullptr && e.parts != &parts)
            throw dcpse_error("expression and target belong to different particle sets");
        field_expr<V> ee = e;
        ee.parts = &parts;
        ee.evaluate_into(prp);
        return *this;
    }
    template <unsigned int prp2>
    field_ref& operator=(const field_ref<prp2, V>& o) {
        return *this = o.as_expr();
    }
    field_ref& operator=(const field_ref& o) { return *this = o.as_expr(); }
};

template <unsigned int prp, typename V>
field_ref<prp, V> getV(V& parts) {
    return field_ref<prp, V>(parts);
}

// ---- expression algebra -------------------------------------------------------------------
template <typename V>
field_expr<V> operator+(field_expr<V> a, const field_expr<V>& b) {
    if (a.parts == nullptr) a.parts = b.parts;
    if (b.parts != nullptr && a.parts != b.parts) throw dcpse_error("terms belong to different particle sets");
    a.ops.insert(a.ops.end(), b.ops.begin(), b.ops.end());
    a.fields.insert(a.fields.end(), b.fields.begin(), b.fields.end());
    return a;
}
template <typename V>
field_expr<V> operator-(const field_expr<V>& a, const field_expr<V>& b) {
    return a + (-b);
}
template <typename V>
field_expr<V> operator*(double s, const field_expr<V>& a) {
    return a.scaled(s);
}
template <typename V>
field_expr<V> operator*(const field_expr<V>& a, double s) {
    return a.scaled(s);
}
template <unsigned int p1, typename V>
field_expr<V> operator*(double s, const field_ref<p1, V>& a) {
    return a.as_expr().scaled(s);
}
template <unsigned int p1, typename V>
field_expr<V> operator*(const field_ref<p1, V>& a, double s) {
    return a.as_expr().scaled(s);
}
template <unsigned int p1, typename V>
field_expr<V> operator+(const field_expr<V>& a, const field_ref<p1, V>& b) {
    return a + b.as_expr();
}
template <unsigned int p1, typename V>
field_expr<V> operator+(const field_ref<p1, V>& a, const field_expr<V>& b) {
    return a.as_expr() + b;
}
template <unsigned int p1, typename V>
field_expr<V> operator-(const field_expr<V>& a, const field_ref<p1, V>& b) {
    return a - b.as_expr();
}
template <unsigned int p1, typename V>
field_expr<V> operator-(const field_ref<p1, V>& a, const field_expr<V>& b) {
    return a.as_expr() - b;
}
template <unsigned int p1, unsigned int p2, typename V>
field_expr<V> operator+(const field_ref<p1, V>& a, const field_ref<p2, V>& b) {
    return a.as_expr() + b.as_expr();
}
template <unsigned int p1, unsigned int p2, typename V>
field_expr<V> operator-(const field_ref<p1, V>& a, const field_ref<p2, V>& b) {
    return a.as_expr() - b.as_expr();
}

// ---- operator classes ------------------------------------------------------------------------
// Field operator over a Dcpse_type that offers only the reference Dcpse
// contract (per-particle computeDifferentialOperator(key, o1)): the whole
// field is evaluated particle by particle on the host, as the reference's
// expression loop does (vector_dist_operators.hpp:1113-1138 over
// VECT_DCPSE / VECT_DCPSE_V_SUM nodes, DCPSE_op.hpp:29-106, 364-442), summing
// the terms' results (the Laplacian's dim operators).
template <typename V, typename D>
class keyed_operator : public field_operator<typename V::prop_type> {
    typedef typename V::prop_type P;

   public:
    struct host_field {
        const std::vector<P>* f;
        V* parts;
        P value(const vect_dist_key_dx& k) const { return (*f)[k.getKey()]; }
        V& getVector() const { return *parts; }
    };
    keyed_operator(V& parts, std::vector<D*> terms) : parts_(parts), terms_(std::move(terms)) {}
    void apply(const P* f, P* out, P alpha, P beta) override {
        const size_t n = parts_.size_local();
        std::vector<P> fh(n), oh(n, P(0));
        if (n) hip_check(hipMemcpy(fh.data(), f, n * sizeof(P), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (n && beta != P(0)) hip_check(hipMemcpy(oh.data(), out, n * sizeof(P), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        host_field o1{&fh, &parts_};
        for (size_t p = 0; p < n; ++p) {
            P acc = P(0);
            for (D* t : terms_) acc += (P)t->computeDifferentialOperator(vect_dist_key_dx(p), o1);
            oh[p] = alpha * acc + (beta != P(0) ? beta * oh[p] : P(0));
        }
        if (n) hip_check(hipMemcpy(out, oh.data(), n * sizeof(P), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }

   private:
    V& parts_;
    std::vector<D*> terms_;
};

// Sum of device field operators (a Laplacian of dim separate Dcpse_type
// objects that do apply whole fields).
template <typename P>
class summed_operator : public field_operator<P> {
   public:
    explicit summed_operator(std::vector<field_operator<P>*> terms) : terms_(std::move(terms)) {}
    void apply(const P* f, P* out, P alpha, P beta) override {
        bool first = true;
        for (auto* t : terms_) {
            t->apply(f, out, alpha, first ? beta : P(1));
            first = false;
        }
    }

   private:
    std::vector<field_operator<P>*> terms_;
};

// Owns the Dcpse_type instance(s) over particle set type V and the field
// operator that expressions apply; forwards the lifecycle calls of the
// reference operator classes.
template <typename VerletList_type, template <unsigned int, typename, typename...> class Dcpse_type>
class dcpse_operator_base {
   protected:
    std::vector<std::shared_ptr<void>> dcpse_;  // one Dcpse_type per term
    std::shared_ptr<void> fop_;                 // field_operator<P> the expressions apply

    template <typename V>
    using dcpse_t = Dcpse_type<V::dims, VerletList_type, V>;

    template <typename V>
    dcpse_t<V>* get(size_t i = 0) const {
        return static_cast<dcpse_t<V>*>(dcpse_.at(i).get());
    }
    template <typename D, typename... Args>
    D* add_term(Args&&... args) {
        D* d = new D(std::forward<Args>(args)...);
        dcpse_.emplace_back(d, [](void* p) { delete static_cast<D*>(p); });
        return d;
    }
    // the field operator over the terms: the Dcpse_type itself when it applies
    // whole fields on the device, else the per-particle host evaluation
    template <typename V, typename D>
    void bind(V& parts) {
        typedef typename V::prop_type P;
        std::vector<D*> terms;
        for (auto& d : dcpse_) terms.push_back(static_cast<D*>(d.get()));
        if constexpr (std::is_base_of<field_operator<P>, D>::value) {
            if (terms.size() == 1) {
                fop_ = std::shared_ptr<void>(dcpse_[0], static_cast<field_operator<P>*>(terms[0]));
            } else {
                std::vector<field_operator<P>*> fo(terms.begin(), terms.end());
                auto* so = new summed_operator<P>(fo);
                fop_ = std::shared_ptr<void>(so, [](void* p) { delete static_cast<summed_operator<P>*>(p); });
            }
        } else {
            auto* ko = new keyed_operator<V, D>(parts, terms);
            fop_ = std::shared_ptr<void>(ko, [](void* p) { delete static_cast<keyed_operator<V, D>*>(p); });
        }
    }
    template <typename V, typename... Args>
    void make(V& parts, Args&&... args) {
        add_term<dcpse_t<V>>(parts, std::forward<Args>(args)...);
        bind<V, dcpse_t<V>>(parts);
    }
    // two sets: support parts, domain subset dom
    template <typename V, typename S, typename... Args>
    void make2(V& parts, S& dom, Args&&... args) {
        typedef Dcpse_type<V::dims, VerletList_type, V, S> D;
        add_term<D>(parts, dom, std::forward<Args>(args)...);
        bind<V, D>(parts);
    }

   public:
    template <unsigned int prp, typename V>
    field_expr<V> operator()(const field_ref<prp, V>& f) const {
        field_expr<V> e;
        e.parts = &f.parts;
        e.ops.push_back({static_cast<field_operator<typename V::prop_type>*>(fop_.get()), prp, 1.0});
        return e;
    }
    template <typename V>
    void update(V& parts) {
        for (size_t i = 0; i < dcpse_.size(); ++i) get<V>(i)->initializeUpdate(parts);
    }
    template <typename V>
    auto checkMomenta(V& parts) {
        return get<V>()->checkMomenta(parts);
    }
    template <typename V>
    void save(V&, const std::string& file) {
        get<V>()->save(file);
    }
    template <typename V>
    void load(V&, const std::string& file) {
        get<V>()->load(file);
    }
    template <unsigned int prp, typename V>
    void DrawKernel(V& parts, int k) {
        get<V>()->template DrawKernel<prp>(parts, k);
    }
    template <unsigned int prp, typename V>
    void DrawKernelNN(V& parts, int k) {
        get<V>()->template DrawKernelNN<prp>(parts, k);
    }
    template <typename V>
    void deallocate(V&) {
        fop_.reset();
        dcpse_.clear();
    }
    // the underlying Dcpse_type object (accessors getNumNN, getCoeffNN, ...);
    // i selects a term of a Laplacian held as dim operators
    template <typename V>
    dcpse_t<V>& dcpse(V&, size_t i = 0) {
        return *get<V>(i);
    }
    // the Dcpse_type of a two-set operator
    template <typename V, typename S>
    Dcpse_type<V::dims, VerletList_type, V, S>& dcpse(V&, S&) {
        return *static_cast<Dcpse_type<V::dims, VerletList_type, V, S>*>(dcpse_.at(0).get());
    }
    size_t terms() const { return dcpse_.size(); }
};

// One operator of a fixed signature (unit exponent vectors and their pairs).
template <typename VerletList_type, template <unsigned int, typename, typename...> class Dcpse_type, unsigned int a0,
          unsigned int a1, unsigned int a2>
class dcpse_signature_operator : public dcpse_operator_base<VerletList_type, Dcpse_type> {
    template <unsigned int dims>
    static Point<dims, unsigned int> signature() {
        static_assert(dims == 3 || a2 == 0, "z derivative of a 2D particle set");
        Point<dims, unsigned int> p;
        p.zero();
        p.get(0) = a0;
        p.get(1) = a1;
        if constexpr (dims == 3) p.get(2) = a2;
        return p;
    }

   public:
    template <typename V>
    dcpse_signature_operator(V& parts, VerletList_type& verletList, unsigned int ord, typename V::stype rCut,
                             support_options opt = support_options::RADIUS) {
        this->template make<V>(parts, verletList, signature<V::dims>(), ord, rCut, opt);
    }
    // support set parts, domain subset dom (DCPSE_op.hpp:693-718)
    template <typename V, typename S, typename = typename S::parent_t>
    dcpse_signature_operator(V& parts, S& dom, VerletList_type& verletList, unsigned int ord, typename V::stype rCut,
                             support_options opt = support_options::RADIUS) {
        this->template make2<V, S>(parts, dom, verletList, signature<V::dims>(), ord, rCut, opt);
    }
};

template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_x_T = dcpse_signature_operator<VL, D, 1, 0, 0>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_y_T = dcpse_signature_operator<VL, D, 0, 1, 0>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_z_T = dcpse_signature_operator<VL, D, 0, 0, 1>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_xx_T = dcpse_signature_operator<VL, D, 2, 0, 0>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_yy_T = dcpse_signature_operator<VL, D, 0, 2, 0>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_zz_T = dcpse_signature_operator<VL, D, 0, 0, 2>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_xy_T = dcpse_signature_operator<VL, D, 1, 1, 0>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_yz_T = dcpse_signature_operator<VL, D, 0, 1, 1>;
template <typename VL, template <unsigned int, typename, typename...> class D = Dcpse_hip>
using Derivative_xz_T = dcpse_signature_operator<VL, D, 1, 0, 1>;

// Arbitrary signature (reference Derivative_G_T, DCPSE_op.hpp:3386-3470).
template <typename VerletList_type, template <unsigned int, typename, typename...> class Dcpse_type = Dcpse_hip>
class Derivative_G_T : public dcpse_operator_base<VerletList_type, Dcpse_type> {
   public:
    template <typename V>
    Derivative_G_T(V& parts, VerletList_type& verletList, unsigned int ord, typename V::stype rCut,
                   const Point<V::dims, unsigned int>& p, support_options opt = support_options::RADIUS) {
        this->template make<V>(parts, verletList, p, ord, rCut, opt);
    }
};

// Laplacian. With a Dcpse_type that takes a combination of signatures
// (Dcpse_hip) it is one operator of signature sum_d 2 e_d; with one that
// offers only the reference constructor it is dim operators of signature
// 2 e_d whose results are summed, as the reference builds it
// (DCPSE_op.hpp:1369-1389, VECT_DCPSE_V_SUM :364-442).
template <typename VerletList_type, template <unsigned int, typename, typename...> class Dcpse_type = Dcpse_hip>
class Laplacian_T : public dcpse_operator_base<VerletList_type, Dcpse_type> {
    template <typename V>
    static constexpr bool fused_ctor =
        std::is_constructible<Dcpse_type<V::dims, VerletList_type, V>, V&, VerletList_type&,
                              const std::vector<uint32_t>&, const std::vector<double>&, unsigned int,
                              typename V::stype, support_options>::value;

   public:
    template <typename V>
    Laplacian_T(V& parts, VerletList_type& verletList, unsigned int ord, typename V::stype rCut,
                support_options opt = support_options::RADIUS) {
        typedef Dcpse_type<V::dims, VerletList_type, V> D;
        if constexpr (fused_ctor<V>) {
            std::vector<uint32_t> sigs(V::dims * V::dims, 0u);
            for (unsigned d = 0; d < V::dims; ++d) sigs[d * V::dims + d] = 2;
            std::vector<double> coeffs(V::dims, 1.0);
            this->template make<V>(parts, verletList, sigs, coeffs, ord, rCut, opt);
        } else {
            for (unsigned i = 0; i < V::dims; ++i) {
                Point<V::dims, unsigned int> p;
                p.zero();
                p.get(i) = 2;
                this->template add_term<D>(parts, verletList, p, ord, rCut, opt);
            }
            this->template bind<V, D>(parts);
        }
    }
};

}  // namespace openfpm_dcpse

template <typename VL>
using Derivative_x = openfpm_dcpse::Derivative_x_T<VL>;
template <typename VL>
using Derivative_y = openfpm_dcpse::Derivative_y_T<VL>;
template <typename VL>
using Derivative_z = openfpm_dcpse::Derivative_z_T<VL>;
template <typename VL>
using Derivative_xx = openfpm_dcpse::Derivative_xx_T<VL>;
template <typename VL>
using Derivative_yy = openfpm_dcpse::Derivative_yy_T<VL>;
template <typename VL>
using Derivative_zz = openfpm_dcpse::Derivative_zz_T<VL>;
template <typename VL>
using Derivative_xy = openfpm_dcpse::Derivative_xy_T<VL>;
template <typename VL>
using Derivative_yz = openfpm_dcpse::Derivative_yz_T<VL>;
template <typename VL>
using Derivative_xz = openfpm_dcpse::Derivative_xz_T<VL>;
template <typename VL>
using Derivative_G = openfpm_dcpse::Derivative_G_T<VL>;
template <typename VL>
using Laplacian = openfpm_dcpse::Laplacian_T<VL>;
using openfpm_dcpse::getV;
