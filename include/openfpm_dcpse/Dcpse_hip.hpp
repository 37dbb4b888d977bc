// Dcpse_hip<dim, VerletList, vector_type>: the Dcpse_type plug point of the
// reference operator classes (template<unsigned, typename, typename...> class
// Dcpse_type, src/DCPSE/DCPSE_op/DCPSE_op.hpp:649) implemented on the gfx950
// kernels of libdcpse_hip.so. It mirrors the reference Dcpse
// (src/DCPSE/Dcpse.hpp:43-826): the same constructors (:91-130), the bulk
// computeDifferentialOperator<f, Df> (:467-499), the row accessors getNumNN /
// getCoeffNN / getIndexNN / getSign / getEpsilonInvPrefactor (:507-550),
// initializeUpdate (:661-687), save / load (:302-379), checkMomenta (:381-457)
// and DrawKernel (:151-193). The per-particle computeDifferentialOperator(key, o1) and
// (key, o1, i) (:560-659) that the reference's expression nodes call
// (VECT_DCPSE, DCPSE_op.hpp:64-67) are served from a host copy of the
// kernels; the operator classes of DCPSE_op_hip.hpp evaluate whole fields
// with device applies instead.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "dcpse_runtime.hpp"
#include "vector_dist_hip.hpp"

#ifndef OPENFPM_DCPSE_NO_VALUE_TYPES
enum support_options { RADIUS, LOAD, ADAPTIVE, AT_LEAST_N_PARTICLES };
#endif

namespace openfpm_dcpse {

// File layout shared with the Python package: "DCPSEHIP", u32 version, dim,
// n_sig, order, f64 h_over_eps, i64 n, nnz, u32 signatures[n_sig*dim],
// f64 coeffs[n_sig], f64 eps[n], eps_inv_pow[n], K[nnz], i64 row_ptr[n+1],
// i32 col[nnz]; little endian, written to <file>_<rank>.
constexpr char kFileMagic[8] = {'D', 'C', 'P', 'S', 'E', 'H', 'I', 'P'};
constexpr uint32_t kFileVersion = 1;

}  // namespace openfpm_dcpse

// The trailing pack matches the reference plug point's
// template<unsigned int, typename, typename...> signature; its one optional
// entry is the domain set type of a two-set operator (a
// vector_dist_subset_hip of the support set), as the reference's
// Dcpse<dim, VerletList, vector_type, vector_type2 = vector_type>.
template <unsigned int dim, typename VerletList_type, typename vector_type, typename... vector_type2>
class Dcpse_hip : public openfpm_dcpse::field_operator<typename vector_type::prop_type> {
    static_assert(sizeof...(vector_type2) <= 1, "one domain set at most");

   public:
    typedef typename vector_type::stype T;
    typedef typename vector_type::prop_type prop_type;
    typedef vector_type vtype;
    typedef std::tuple_element_t<0, std::tuple<vector_type2..., vector_type>> domain_type;
    static constexpr bool two_sets = sizeof...(vector_type2) == 1;
    // Single-field applies run the operator's sliced-ELL copy (FP32: 16-bit
    // column offsets, FP64: 32-bit columns) from this apply since assembly
    // on, the CSR kernel before. 1: from the first, so every apply of an
    // operator runs one kernel and repeated applies agree bit for bit.
    static constexpr int kSellAfter = 1;

    T HOverEpsilon = 0.9;
    // one step of FP64 iterative refinement of each particle's solve
    // (DCPSE_ASM_REFINE); off as in the reference, whose QR solve has none
    bool refine = false;

    template <typename PointT>
    Dcpse_hip(vector_type& particles, VerletList_type& verletList, PointT differentialSignature,
              unsigned int convergenceOrder, T rCut, support_options opt = support_options::RADIUS)
        : particles_(particles), verletList_(verletList), convergenceOrder_(convergenceOrder), rCut_(rCut) {
        std::vector<uint32_t> sig(dim);
        for (unsigned d = 0; d < dim; ++d) sig[d] = (uint32_t)differentialSignature.value(d);
        sigs_ = sig;
        coeffs_ = {1.0};
        finish_init(opt);
    }

    // Two sets (Dcpse.hpp:111-130): kernels of the domain particles over
    // neighbours in the support set; the domain is a subset of the support
    // set (vector_dist_subset_hip) and the Verlet list its getVerlet().
    template <typename PointT, bool B = two_sets, typename = std::enable_if_t<B>>
    Dcpse_hip(vector_type& particlesSupport, domain_type& particlesDomain, VerletList_type& verletList,
              PointT differentialSignature, unsigned int convergenceOrder, T rCut,
              support_options opt = support_options::RADIUS)
        : particles_(particlesSupport), verletList_(verletList), convergenceOrder_(convergenceOrder), rCut_(rCut),
          domain_(&particlesDomain) {
        std::vector<uint32_t> sig(dim);
        for (unsigned d = 0; d < dim; ++d) sig[d] = (uint32_t)differentialSignature.value(d);
        sigs_ = sig;
        coeffs_ = {1.0};
        finish_init(opt);
    }

    // Linear combination sum_k coeffs[k] D^{sig_k} of signatures sharing one basis
    // (the Laplacian assembles sum_d D^{2 e_d} this way).
    Dcpse_hip(vector_type& particles, VerletList_type& verletList, const std::vector<uint32_t>& signatures,
              const std::vector<double>& coeffs, unsigned int convergenceOrder, T rCut,
              support_options opt = support_options::RADIUS)
        : particles_(particles), verletList_(verletList), convergenceOrder_(convergenceOrder), rCut_(rCut),
          sigs_(signatures), coeffs_(coeffs) {
        finish_init(opt);
    }

    // ---- apply ---------------------------------------------------------------------------
    // Df = D f over the whole set (Dcpse.hpp:467-499). With two sets, f is read
    // from the support set and Df written at the domain's particles.
    template <unsigned int fValuePos, unsigned int DfValuePos>
    void computeDifferentialOperator(vector_type& particles) {
        if constexpr (two_sets) {
            const size_t nd = domain_->size_local();
            std::vector<prop_type> self(nd);
            for (size_t p = 0; p < nd; ++p)
                self[p] = particles.template getProp<fValuePos>(domain_->supportIndex(p));
            openfpm_dcpse::device_array<prop_type> fs, out(nd);
            fs.upload(self);
            apply_two_set(particles.devProp(fValuePos), fs.data(), out.data(), prop_type(1), prop_type(0));
            const std::vector<prop_type> h = out.to_host();
            for (size_t p = 0; p < nd; ++p) domain_->template getProp<DfValuePos>(p) = h[p];
        } else {
            const prop_type* f = particles.devProp(fValuePos);
            prop_type* out = particles.devProp(DfValuePos);
            apply(f, out, prop_type(1), prop_type(0));
            particles.devPropWritten(DfValuePos);
        }
    }

    // Two sets: out[p] = alpha * eps_p^-|b| sum_q K_pq (f[q] - f_self[p]) + beta * out[p]
    // over the domain rows, f a support-set field, f_self its value at each
    // domain particle (device arrays).
    void apply_two_set(const prop_type* f, const prop_type* f_self, prop_type* out, prop_type alpha,
                       prop_type beta) {
        auto& c = openfpm_dcpse::context();
        check_fresh();
        const int64_t n = (int64_t)verletList_.size(), n_sup = (int64_t)particles_.size_local();
        prop_type* O[1] = {out};
        if constexpr (std::is_same<prop_type, float>::value) {
            const float* W[1] = {W32_.data()};
            c.check(dcpse_apply_f32(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W, f, n_sup, 1,
                                    f_self, n, O, n, alpha, beta),
                    "dcpse_apply_f32");
        } else {
            const double* W[1] = {w64()};
            c.check(dcpse_apply_f64(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W, f, n_sup, 1,
                                    f_self, n, O, n, alpha, beta),
                    "dcpse_apply_f64");
        }
    }

    // out = alpha * D f + beta * out on device arrays of the particle count
    // (two sets: f over the support set, out over the domain; the values at
    // the domain particles are gathered from f on the host)
    void apply(const prop_type* f, prop_type* out, prop_type alpha, prop_type beta) override {
        auto& c = openfpm_dcpse::context();
        check_fresh();
        const int64_t n = (int64_t)verletList_.size();
        if (n == 0) return;  // an empty domain: nothing to compute
        if constexpr (two_sets) {
            std::vector<prop_type> fh(particles_.size_local()), self(domain_->size_local());
            if (!fh.empty())
                openfpm_dcpse::hip_check(hipMemcpy(fh.data(), f, fh.size() * sizeof(prop_type), hipMemcpyDeviceToHost),
                                         "hipMemcpy D2H");
            for (size_t p = 0; p < self.size(); ++p) self[p] = fh[domain_->supportIndex(p)];
            openfpm_dcpse::device_array<prop_type> fs;
            fs.upload(self);
            apply_two_set(f, fs.data(), out, alpha, beta);
            return;
        }
        if constexpr (std::is_same<prop_type, float>::value) {
            float* O[1] = {out};
            // the sliced-ELL copy from the kSellAfter-th apply since assembly
            // on (building it costs what some 50 applies save)
            if (n > 0 && !b16_tried_ && ++n_apply32_ >= kSellAfter) build_bundle16();
            if (b16_) {
                c.check(dcpse_apply_sell_bundle16_f32(c.get(), n, verletList_.devRowPtr(),
                                                      verletList_.sell(false).slice_ptr.data(),
                                                      row16_.data(), 1, bundle_.data(), exc_ptr_.data(),
                                                      exc_col_.data(), exc_W_.data(), n_exc_, f, nullptr, O, alpha,
                                                      beta),
                        "dcpse_apply_sell_bundle16_f32");
                return;
            }
            const float* W[1] = {W32_.data()};
            c.check(dcpse_apply_f32(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W, f, n, 1,
                                    nullptr, n, O, n, alpha, beta),
                    "dcpse_apply_f32");
        } else {
            double* O[1] = {out};
            if (n > 0 && !s64_tried_ && ++n_apply64_ >= kSellAfter) build_sell64();
            if (b64_) {
                c.check(dcpse_apply_sell_bundle16_f64(c.get(), n, verletList_.devRowPtr(),
                                                      verletList_.sell(false).slice_ptr.data(), row16_64_.data(), 1,
                                                      bundle64_.data(), exc_ptr64_.data(), exc_col64_.data(),
                                                      exc_W64_.data(), n_exc64_, f, nullptr, O, alpha, beta),
                        "dcpse_apply_sell_bundle16_f64");
                return;
            }
            if (s64_) {
                const auto& L = verletList_.sell(true);
                const double* SW[1] = {sellW64_.data()};
                c.check(dcpse_apply_sell_f64(c.get(), n, verletList_.devRowPtr(), L.slice_ptr.data(),
                                             L.sell_row.data(), L.sell_col.data(), 1, SW, f, nullptr, O, alpha, beta),
                        "dcpse_apply_sell_f64");
                return;
            }
            const double* W[1] = {w64()};
            c.check(dcpse_apply_f64(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W, f, n, 1,
                                    nullptr, n, O, n, alpha, beta),
                    "dcpse_apply_f64");
        }
    }

    // ---- row accessors (value_nz contract, DCPSE_op.hpp:69-83) ------------------------------
    int getNumNN(const vect_dist_key_dx& p) { return (int)verletList_.getNNPart(p); }
    T getCoeffNN(const vect_dist_key_dx& p, int j) {
        mirror();
        return (T)K_host_[rp_host_[p.getKey()] + j];
    }
    size_t getIndexNN(const vect_dist_key_dx& p, int j) { return verletList_.get(p, j); }
    T getSign() {
        const unsigned o = order_of_derivative();
        return (o % 2 == 0 && o != 0) ? T(-1) : T(1);
    }
    T getEpsilonInvPrefactor(const vect_dist_key_dx& p) {
        mirror();
        return (T)eip_host_[p.getKey()];
    }

    // ---- per-particle apply (Dcpse.hpp:560-659) -------------------------------------------
    // D o1 at domain particle p, o1 an expression with value(key) over the
    // support set (and getVector()): eps_p^-|b| sum_q (o1(q) + s o1(p)) K_pq,
    // s = -1 for even |b|. With two sets o1(p) is read at p's index in the
    // support set. Served from a host copy of the kernels (the reference's
    // host expression loop); whole fields go through the device applies.
    template <typename op_type>
    auto computeDifferentialOperator(const vect_dist_key_dx& p, op_type& o1)
        -> std::decay_t<decltype(o1.value(p))> {
        typedef std::decay_t<decltype(o1.value(p))> expr_type;
        check_fresh();
        mirror();
        const size_t k = p.getKey();
        const expr_type fxp = apply_sign() * o1.value(vect_dist_key_dx(self_index(k)));
        expr_type Dfxp = expr_type(0);
        for (int64_t j = rp_host_[k]; j < rp_host_[k + 1]; ++j)
            Dfxp = Dfxp + (o1.value(vect_dist_key_dx((size_t)col_host_[j])) + fxp) * (T)K_host_[j];
        return Dfxp * (T)eip_host_[k];
    }
    // component i of a vector-valued o1
    template <typename op_type>
    auto computeDifferentialOperator(const vect_dist_key_dx& p, op_type& o1, int i)
        -> std::decay_t<decltype(o1.value(p)[i])> {
        typedef std::decay_t<decltype(o1.value(p)[i])> expr_type;
        check_fresh();
        mirror();
        const size_t k = p.getKey();
        const expr_type fxp = apply_sign() * o1.value(vect_dist_key_dx(self_index(k)))[i];
        expr_type Dfxp = expr_type(0);
        for (int64_t j = rp_host_[k]; j < rp_host_[k + 1]; ++j)
            Dfxp = Dfxp + (o1.value(vect_dist_key_dx((size_t)col_host_[j]))[i] + fxp) * (T)K_host_[j];
        return Dfxp * (T)eip_host_[k];
    }

    // ---- lifecycle ------------------------------------------------------------------------
    // Reassembles on the current positions (Dcpse.hpp:661-687); call after the
    // particles moved and the Verlet list was rebuilt.
    void initializeUpdate(vector_type& particles) {
        (void)particles;
        assemble();
    }
    template <bool B = two_sets, typename = std::enable_if_t<B>>
    void initializeUpdate(vector_type& particlesSupport, domain_type& particlesDomain) {
        (void)particlesSupport;
        domain_ = &particlesDomain;
        assemble();
    }

    template <unsigned int prp>
    void DrawKernel(vector_type& particles, int p) {
        mirror();
        for (int64_t j = rp_host_[p]; j < rp_host_[p + 1]; ++j)
            particles.template getProp<prp>((size_t)col_host_[j]) += (prop_type)K_host_[j];
    }
    template <unsigned int prp>
    void DrawKernelNN(vector_type& particles, int p) {
        mirror();
        for (int64_t j = rp_host_[p]; j < rp_host_[p + 1]; ++j)
            particles.template getProp<prp>((size_t)col_host_[j]) = prop_type(1);
    }

    // Per basis monomial, the smallest and largest over particles of the moment
    // sum_q K_pq m_alpha((x_p - x_q) / eps_p), printed one "MOMENTA:" line per
    // monomial and returned (Dcpse.hpp:381-457). A valid operator gives
    // b_alpha (DcpseRhs.hpp:27-60) at every particle. Periodic neighbours are
    // taken at their nearest image, as the support build found them; both
    // extremes are of the sum itself (the reference's maximum keeps the sum
    // only when an unset slot exceeds it).
    std::vector<std::pair<double, double>> checkMomenta(vector_type& particles) {
        mirror();
        const uint32_t* beta = sigs_.data();
        const int32_t m = dcpse_basis_size(dim, beta, convergenceOrder_);
        std::vector<uint32_t> ex((size_t)std::max(m, 1) * dim);
        if (m < 0 || dcpse_basis(dim, beta, convergenceOrder_, ex.data(), m) != m)
            throw openfpm_dcpse::dcpse_error("checkMomenta: no monomial basis");
        std::vector<std::pair<double, double>> mm((size_t)m, {HUGE_VAL, -HUGE_VAL});
        std::vector<double> acc((size_t)m);
        double len[dim];
        for (unsigned d = 0; d < dim; ++d)
            len[d] = (double)particles.boxHigh()[d] - (double)particles.boxLow()[d];
        const unsigned per = particles.periodicMask();
        const size_t n = eps_host_.size();
        for (size_t p = 0; p < n; ++p) {
            std::fill(acc.begin(), acc.end(), 0.0);
            const auto xp = dom_pos(particles, p);
            for (int64_t j = rp_host_[p]; j < rp_host_[p + 1]; ++j) {
                const auto xq = particles.getPos(vect_dist_key_dx((size_t)col_host_[j]));
                double z[dim];
                for (unsigned d = 0; d < dim; ++d) {
                    double dd = (double)xp[d] - (double)xq[d];
                    if ((per >> d) & 1u) dd -= len[d] * std::nearbyint(dd / len[d]);
                    z[d] = dd / eps_host_[p];
                }
                for (int32_t i = 0; i < m; ++i) {
                    double v = K_host_[j];
                    for (unsigned d = 0; d < dim; ++d)
                        for (uint32_t e = 0; e < ex[(size_t)i * dim + d]; ++e) v *= z[d];
                    acc[(size_t)i] += v;
                }
            }
            for (int32_t i = 0; i < m; ++i) {
                mm[(size_t)i].first = std::min(mm[(size_t)i].first, acc[(size_t)i]);
                mm[(size_t)i].second = std::max(mm[(size_t)i].second, acc[(size_t)i]);
            }
        }
        for (int32_t i = 0; i < m; ++i) {
            std::cout << "MOMENTA: ";
            for (unsigned d = 0; d < dim; ++d) std::cout << (d ? " " : "(") << ex[(size_t)i * dim + d];
            std::cout << ")Min: " << mm[(size_t)i].first << "  Max: " << mm[(size_t)i].second << std::endl;
        }
        return mm;
    }

    // The reference's DCPSE_VERBOSE statistics (Dcpse.hpp:786-795) with the
    // largest moment residual and condition estimate of the solved systems
    // (dcpse_operator_diagnostics); print = true writes the reference's line.
    struct diagnostics_t {
        double n_min, n_mean, n_max, h_mean, min_spacing_mean, min_spacing_min, min_spacing_max, residual_max,
            cond1_max, not_factored;
    };
    diagnostics_t diagnostics(bool print = false) {
        auto& c = openfpm_dcpse::context();
        check_fresh();
        double lo[dim], hi[dim];
        for (unsigned d = 0; d < dim; ++d) {
            lo[d] = (double)particles_.boxLow()[d];
            hi[d] = (double)particles_.boxHigh()[d];
        }
        const double* pdom = nullptr;
        if constexpr (two_sets) pdom = domain_->devPos();
        double out[10] = {0};
        c.check(dcpse_operator_diagnostics(c.get(), dim, particles_.devPos(), (int64_t)particles_.size_local(), pdom,
                                           (int64_t)verletList_.size(), lo, hi, particles_.periodicMask(),
                                           verletList_.devRowPtr(), verletList_.devCol(), (int32_t)coeffs_.size(),
                                           sigs_.data(), coeffs_.data(), convergenceOrder_, (double)HOverEpsilon,
                                           eps_.data(), K64_.size() ? K64_.data() : eps_.data(), nullptr, nullptr,
                                           nullptr, out),
                "dcpse_operator_diagnostics");
        diagnostics_t r{out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7], out[8], out[9]};
        if (print)
            std::cout << "DCPSE Operator Construction Complete. The global avg spacing in the support <h> is: "
                      << r.h_mean << " (c=" << HOverEpsilon << "). Avg:" << r.min_spacing_mean << " Range:["
                      << r.min_spacing_min << "," << r.min_spacing_max << "]. Neighbours " << r.n_min << ".."
                      << r.n_max << " (mean " << r.n_mean << "), moment residual " << r.residual_max
                      << ", cond_1(A) up to " << r.cond1_max << std::endl;
        return r;
    }

    void save(const std::string& file) {
        mirror();
        std::ofstream out(file + "_0", std::ios::binary);
        if (!out) throw openfpm_dcpse::dcpse_error("cannot write " + file + "_0");
        const uint32_t hdr[4] = {openfpm_dcpse::kFileVersion, dim, (uint32_t)coeffs_.size(), convergenceOrder_};
        const double hoe = (double)HOverEpsilon;
        const int64_t n = (int64_t)eps_host_.size(), nnz = (int64_t)K_host_.size();
        out.write(openfpm_dcpse::kFileMagic, 8);
        out.write((const char*)hdr, sizeof(hdr));
        out.write((const char*)&hoe, 8);
        out.write((const char*)&n, 8);
        out.write((const char*)&nnz, 8);
        out.write((const char*)sigs_.data(), 4 * sigs_.size());
        out.write((const char*)coeffs_.data(), 8 * coeffs_.size());
        out.write((const char*)eps_host_.data(), 8 * n);
        out.write((const char*)eip_host_.data(), 8 * n);
        out.write((const char*)K_host_.data(), 8 * nnz);
        out.write((const char*)rp_host_.data(), 8 * (n + 1));
        out.write((const char*)col_host_.data(), 4 * nnz);
    }

    // Loads kernels saved by save() (or by the Python package) for the same
    // signature and support; the Verlet list of this operator must match.
    void load(const std::string& file) {
        std::ifstream in(file + "_0", std::ios::binary);
        if (!in) throw openfpm_dcpse::dcpse_error("cannot open " + file + "_0");
        char magic[8];
        uint32_t hdr[4];
        double hoe;
        int64_t n, nnz;
        in.read(magic, 8);
        in.read((char*)hdr, sizeof(hdr));
        in.read((char*)&hoe, 8);
        in.read((char*)&n, 8);
        in.read((char*)&nnz, 8);
        if (!in || std::memcmp(magic, openfpm_dcpse::kFileMagic, 8) != 0 || hdr[0] != openfpm_dcpse::kFileVersion ||
            hdr[1] != dim)
            throw openfpm_dcpse::dcpse_error(file + "_0 is not a DC-PSE kernel file of this dimension");
        std::vector<uint32_t> sig(hdr[2] * dim);
        std::vector<double> co(hdr[2]);
        in.read((char*)sig.data(), 4 * sig.size());
        in.read((char*)co.data(), 8 * co.size());
        if (sig != sigs_) throw openfpm_dcpse::dcpse_error("kernel file holds a different differential signature");
        if ((size_t)n != verletList_.size() || (size_t)nnz != verletList_.nnz())
            throw openfpm_dcpse::dcpse_error("kernel file support differs from this operator's Verlet list");
        eps_host_.resize(n);
        eip_host_.resize(n);
        K_host_.resize(nnz);
        rp_host_.resize(n + 1);
        col_host_.resize(nnz);
        in.read((char*)eps_host_.data(), 8 * n);
        in.read((char*)eip_host_.data(), 8 * n);
        in.read((char*)K_host_.data(), 8 * nnz);
        in.read((char*)rp_host_.data(), 8 * (n + 1));
        in.read((char*)col_host_.data(), 4 * nnz);
        if (!in) throw openfpm_dcpse::dcpse_error(file + "_0 is truncated");
        if (rp_host_ != verletList_.hostRowPtr() || col_host_ != verletList_.hostCol())
            throw openfpm_dcpse::dcpse_error("kernel file support differs from this operator's Verlet list");
        convergenceOrder_ = hdr[3];
        HOverEpsilon = (T)hoe;
        coeffs_ = co;
        // pre-scaled weights W = eps^-|beta| K for the apply kernels
        std::vector<double> W64(nnz);
        std::vector<float> W32(nnz);
        for (int64_t p = 0; p < n; ++p)
            for (int64_t j = rp_host_[p]; j < rp_host_[p + 1]; ++j) {
                W64[j] = K_host_[j] * eip_host_[p];
                W32[j] = (float)W64[j];
            }
        K64_.upload(K_host_);
        W64_.upload(W64);
        W32_.upload(W32);
        b16_ = b16_tried_ = false;
        n_apply32_ = 0;
        s64_ = s64_tried_ = b64_ = false;
        n_apply64_ = 0;
        mirrored_ = true;
        ctr_ = particles_.getMapCtr();
    }

    const std::vector<uint32_t>& signatures() const { return sigs_; }

    // FP32 applies run on a sliced-ELL copy of the weights with 16-bit column
    // offsets (dcpse_sell_bundle16_count/fill over the Verlet list's shared
    // slice structure), built at the kSellAfter-th apply after each assembly
    // or load. Where it cannot be built (no device memory for it, or a row
    // count beyond the 16-bit layout's range) the CSR kernel runs instead.
    void build_bundle16() {
        auto& c = openfpm_dcpse::context();
        b16_tried_ = true;
        b16_ = false;
        try {
            const auto& L = verletList_.sell(false);
            const int64_t n = (int64_t)verletList_.size();
            const int64_t n_lanes = 64 * ((n + 63) / 64);
            const int64_t n_slots = L.n_slots;
            int64_t n_exc = 0;
            exc_ptr_.resize(n_lanes + 1);
            if (dcpse_sell_bundle16_count(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(),
                                          L.sell_row.data(), exc_ptr_.data(), &n_exc))
                return release_bundle16();
            bundle_.resize((size_t)std::max<int64_t>(n_slots / 2 + n_slots, 4));
            row16_.resize(n_lanes);
            exc_col_.resize((size_t)std::max<int64_t>(n_exc, 1));
            exc_W_.resize((size_t)std::max<int64_t>(n_exc, 1));
            const float* W[1] = {W32_.data()};
            if (dcpse_sell_fill_bundle16_f32(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W,
                                             L.slice_ptr.data(), L.sell_row.data(), exc_ptr_.data(), n_exc,
                                             bundle_.data(), row16_.data(), exc_col_.data(), exc_W_.data()))
                return release_bundle16();
            n_exc_ = n_exc;
            b16_ = true;
        } catch (const openfpm_dcpse::dcpse_error&) {
            release_bundle16();
        }
    }

    // FP64 applies run on a sliced-ELL copy over the Verlet list's shared
    // slices, built at the kSellAfter-th FP64 apply after each assembly or
    // load: the 16-bit-offset FP64 bundle (one set; 18 bytes per neighbour),
    // else FP64 weights in the slot order of the 32-bit-column layout, else
    // (no device memory for either) the CSR kernel.
    bool sell64_built() const { return s64_ || b64_; }
    void build_sell64() {
        auto& c = openfpm_dcpse::context();
        s64_tried_ = true;
        s64_ = false;
        if constexpr (!two_sets) {
            if (build_bundle16_64()) return;
        }
        try {
            const auto& L = verletList_.sell(true);
            const int64_t n = (int64_t)verletList_.size();
            sellW64_.resize((size_t)std::max<int64_t>(L.n_slots, 4));
            if (dcpse_sell_fill_weights_f64(c.get(), n, verletList_.devRowPtr(), w64(), L.slice_ptr.data(),
                                            L.sell_row.data(), sellW64_.data())) {
                sellW64_ = openfpm_dcpse::device_array<double>();
                return;
            }
            s64_ = true;
        } catch (const openfpm_dcpse::dcpse_error&) {
            sellW64_ = openfpm_dcpse::device_array<double>();
        }
    }

   private:
    // FP64 pre-scaled weights eps^-|beta| K, formed from the kernels at the
    // first FP64 use (the assembly writes only K and the FP32 weights)
    const double* w64() {
        const size_t nnz = verletList_.nnz();
        if (W64_.size() != nnz) {
            auto& c = openfpm_dcpse::context();
            W64_.resize(nnz);
            c.check(dcpse_scale_rows_f64(c.get(), (int64_t)verletList_.size(), verletList_.devRowPtr(), K64_.data(),
                                         eip_.data(), W64_.data()),
                    "dcpse_scale_rows_f64");
        }
        return W64_.data();
    }
    bool build_bundle16_64() {
        auto& c = openfpm_dcpse::context();
        b64_ = false;
        try {
            const auto& L = verletList_.sell(false);
            const int64_t n = (int64_t)verletList_.size();
            const int64_t n_lanes = 64 * ((n + 63) / 64);
            const int64_t n_slots = L.n_slots;
            int64_t n_exc = 0;
            exc_ptr64_.resize(n_lanes + 1);
            if (dcpse_sell_bundle16_count(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(),
                                          L.sell_row.data(), exc_ptr64_.data(), &n_exc)) {
                release_bundle16_64();
                return false;
            }
            bundle64_.resize((size_t)std::max<int64_t>(n_slots / 2 + 2 * n_slots, 8));
            row16_64_.resize(n_lanes);
            exc_col64_.resize((size_t)std::max<int64_t>(n_exc, 1));
            exc_W64_.resize((size_t)std::max<int64_t>(n_exc, 1));
            const double* W[1] = {w64()};
            if (dcpse_sell_fill_bundle16_f64(c.get(), n, verletList_.devRowPtr(), verletList_.devCol(), 1, W,
                                             L.slice_ptr.data(), L.sell_row.data(), exc_ptr64_.data(), n_exc,
                                             bundle64_.data(), row16_64_.data(), exc_col64_.data(),
                                             exc_W64_.data())) {
                release_bundle16_64();
                return false;
            }
            n_exc64_ = n_exc;
            b64_ = true;
        } catch (const openfpm_dcpse::dcpse_error&) {
            release_bundle16_64();
        }
        return b64_;
    }
    void release_bundle16_64() {
        b64_ = false;
        exc_ptr64_ = openfpm_dcpse::device_array<int64_t>();
        row16_64_ = openfpm_dcpse::device_array<int32_t>();
        bundle64_ = openfpm_dcpse::device_array<int32_t>();
        exc_col64_ = openfpm_dcpse::device_array<int32_t>();
        exc_W64_ = openfpm_dcpse::device_array<double>();
    }
    void release_bundle16() {
        b16_ = false;
        exc_ptr_ = openfpm_dcpse::device_array<int64_t>();
        row16_ = openfpm_dcpse::device_array<int32_t>();
        bundle_ = openfpm_dcpse::device_array<int32_t>();
        exc_col_ = openfpm_dcpse::device_array<int32_t>();
        exc_W_ = openfpm_dcpse::device_array<float>();
    }
    // the reference apply's s: -1 for even |beta| (zero included), +1 for odd
    T apply_sign() const { return order_of_derivative() % 2 == 0 ? T(-1) : T(1); }
    size_t self_index(size_t p) const {
        if constexpr (two_sets) return domain_->supportIndex(p);
        return p;
    }
    Point<dim, T> dom_pos(vector_type& particles, size_t p) const {
        if constexpr (two_sets) return domain_->getPos(vect_dist_key_dx(p));
        return particles.getPos(vect_dist_key_dx(p));
    }
    unsigned order_of_derivative() const {
        unsigned o = 0;
        for (unsigned d = 0; d < dim; ++d) o += sigs_[d];
        return o;
    }
    void finish_init(support_options opt) {
        if (opt != support_options::LOAD) assemble();
    }
    void check_fresh() {
        if (K64_.size() == 0 && verletList_.nnz() != 0)
            throw openfpm_dcpse::dcpse_error("DC-PSE operator has no kernels: call load() or update()");
        if (ctr_ != particles_.getMapCtr())
            throw openfpm_dcpse::dcpse_error("particles moved since assembly: call update() first");
    }
    void assemble() {
        auto& c = openfpm_dcpse::context();
        const size_t n = verletList_.size(), nnz = verletList_.nnz();
        eps_.resize(n);
        eip_.resize(n);
        info_.resize(n);
        K64_.resize(nnz);
        W64_ = openfpm_dcpse::device_array<double>();  // derived at the first FP64 use (w64)
        W32_.resize(nnz);
        if (n == 0) {  // an empty domain: no rows to assemble
            mirrored_ = false;
            ctr_ = particles_.getMapCtr();
            return;
        }
        double lo[dim], hi[dim];
        for (unsigned d = 0; d < dim; ++d) {
            lo[d] = (double)particles_.boxLow()[d];
            hi[d] = (double)particles_.boxHigh()[d];
        }
        const double* pdom = nullptr;
        if constexpr (two_sets) pdom = domain_->devPos();
        c.check(dcpse_assemble(c.get(), dim, particles_.devPos(), (int64_t)particles_.size_local(), pdom, (int64_t)n,
                               lo, hi,
                               particles_.periodicMask(), verletList_.devRowPtr(), verletList_.devCol(),
                               (int32_t)coeffs_.size(), sigs_.data(), coeffs_.data(), convergenceOrder_,
                               (double)HOverEpsilon, refine ? DCPSE_ASM_REFINE : 0u, eps_.data(), eip_.data(), K64_.data(),
                               W32_.data(), nullptr, info_.data()),
                "dcpse_assemble");
        const std::vector<int32_t> info = info_.to_host();
        for (size_t p = 0; p < n; ++p)
            if (info[p] == DCPSE_INFO_TOO_FEW)
                throw std::length_error("Not enough neighbour points passed for Vandermonde matrix construction!");
            else if (info[p] != DCPSE_INFO_OK)
                throw openfpm_dcpse::dcpse_error("DC-PSE assembly failed at particle " + std::to_string(p) +
                                                 " (info " + std::to_string(info[p]) + ")");
        mirrored_ = false;
        b16_ = b16_tried_ = false;
        n_apply32_ = 0;
        s64_ = s64_tried_ = b64_ = false;
        n_apply64_ = 0;
        ctr_ = particles_.getMapCtr();
    }
    void mirror() {
        if (mirrored_) return;
        eps_host_ = eps_.to_host();
        eip_host_ = eip_.to_host();
        K_host_ = K64_.to_host();
        std::vector<int64_t> rp(verletList_.size() + 1);
        hip_copy(rp.data(), verletList_.devRowPtr(), rp.size());
        rp_host_ = rp;
        col_host_.resize(verletList_.nnz());
        hip_copy(col_host_.data(), verletList_.devCol(), col_host_.size());
        mirrored_ = true;
    }
    template <typename U>
    static void hip_copy(U* dst, const U* src, size_t n) {
        if (n) openfpm_dcpse::hip_check(hipMemcpy(dst, src, n * sizeof(U), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }

    vector_type& particles_;
    VerletList_type& verletList_;
    unsigned int convergenceOrder_;
    T rCut_;
    domain_type* domain_ = nullptr;  // two sets: the domain subset
    std::vector<uint32_t> sigs_;
    std::vector<double> coeffs_;
    openfpm_dcpse::device_array<double> eps_, eip_, K64_, W64_;
    openfpm_dcpse::device_array<float> W32_;
    openfpm_dcpse::device_array<int32_t> info_;
    // the 16-bit sliced-ELL copy used by FP32 applies (over the list's slices)
    openfpm_dcpse::device_array<int64_t> exc_ptr_;
    openfpm_dcpse::device_array<int32_t> row16_, bundle_, exc_col_;
    openfpm_dcpse::device_array<float> exc_W_;
    int64_t n_exc_ = 0;
    bool b16_ = false, b16_tried_ = false;
    int n_apply32_ = 0;  // FP32 applies since assembly
    // FP64 weights in the list's sliced-ELL slot order, used by FP64 applies
    openfpm_dcpse::device_array<double> sellW64_;
    bool s64_ = false, s64_tried_ = false;
    // the 16-bit FP64 bundle used by FP64 applies of one set (over the list's slices)
    openfpm_dcpse::device_array<int64_t> exc_ptr64_;
    openfpm_dcpse::device_array<int32_t> row16_64_, bundle64_, exc_col64_;
    openfpm_dcpse::device_array<double> exc_W64_;
    int64_t n_exc64_ = 0;
    bool b64_ = false;
    int n_apply64_ = 0;  // FP64 applies since assembly
    std::vector<double> eps_host_, eip_host_, K_host_;
    std::vector<int64_t> rp_host_;
    std::vector<int32_t> col_host_;
    bool mirrored_ = false;
    size_t ctr_ = 0;
};
