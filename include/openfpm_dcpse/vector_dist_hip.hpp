// A particle container and Verlet list with the parts of openfpm's
// vector_dist / VerletList interface the DC-PSE path uses (add, getLastPos,
// getLastProp, getPos, getProp, getDomainIterator, getVerlet, map, ghost_get,
// size_local), backed by device memory for the HIP kernels. One GPU, one
// process: map() and ghost_get() have nothing to exchange; periodic
// boundaries are handled by minimum-image displacements instead of ghosts.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <type_traits>
#include <vector>

#include "dcpse_runtime.hpp"

// ---- minimal openfpm-style value types -------------------------------------------
#ifndef OPENFPM_DCPSE_NO_VALUE_TYPES
template <unsigned int dim, typename T>
class Point {
   public:
    Point() { zero(); }
    Point(std::initializer_list<T> l) {
        zero();
        unsigned i = 0;
        for (T v : l)
            if (i < dim) x[i++] = v;
    }
    void zero() {
        for (unsigned i = 0; i < dim; ++i) x[i] = T(0);
    }
    T& get(unsigned i) { return x[i]; }
    T value(unsigned i) const { return x[i]; }
    T operator[](unsigned i) const { return x[i]; }
    T& operator[](unsigned i) { return x[i]; }
    const T* asArray() const { return x; }

   private:
    T x[dim];
};

template <unsigned int dim, typename T>
class Box {
   public:
    Box(std::initializer_list<T> lo, std::initializer_list<T> hi) {
        unsigned i = 0;
        for (T v : lo) lo_[i++] = v;
        i = 0;
        for (T v : hi) hi_[i++] = v;
    }
    T getLow(unsigned i) const { return lo_[i]; }
    T getHigh(unsigned i) const { return hi_[i]; }

   private:
    T lo_[dim], hi_[dim];
};

template <unsigned int dim, typename T>
struct Ghost {
    explicit Ghost(T) {}
};

constexpr size_t NON_PERIODIC = 0;
constexpr size_t PERIODIC = 1;
constexpr unsigned int VL_NON_SYMMETRIC = 1;
constexpr unsigned int VL_SKIP_REF_PART = 2;

typedef long int mem_id;

struct vect_dist_key_dx {
    size_t k = 0;
    vect_dist_key_dx() = default;
    explicit vect_dist_key_dx(size_t key) : k(key) {}
    size_t getKey() const { return k; }
};
#endif

namespace openfpm_dcpse {

// Lexicographic iterator over an index grid (dimension 0 fastest).
template <unsigned int dim>
class grid_key_iterator {
   public:
    struct key {
        size_t k[dim];
        long get(unsigned d) const { return (long)k[d]; }
    };
    explicit grid_key_iterator(const size_t (&sz)[dim]) {
        for (unsigned d = 0; d < dim; ++d) {
            sz_[d] = sz[d];
            cur_.k[d] = 0;
            if (sz[d] == 0) done_ = true;
        }
    }
    bool isNext() const { return !done_; }
    key get() const { return cur_; }
    grid_key_iterator& operator++() {
        for (unsigned d = 0; d < dim; ++d) {
            if (++cur_.k[d] < sz_[d]) return *this;
            cur_.k[d] = 0;
        }
        done_ = true;
        return *this;
    }

   private:
    size_t sz_[dim];
    key cur_;
    bool done_ = false;
};

template <typename vector_type>
class VerletList_hip;
template <typename parent_type>
class vector_dist_subset_hip;

// Scalar properties only (the DC-PSE operators act on scalar fields); nprop of
// them, of type P (double or float: float fields run the FP32 kernels).
template <unsigned int dim, typename St = double, unsigned int nprop = 3, typename P = St>
class vector_dist_hip {
   public:
    static constexpr unsigned int dims = dim;
    typedef St stype;
    typedef P prop_type;

    vector_dist_hip(size_t n, const Box<dim, St>& box, const size_t (&bc)[dim], const Ghost<dim, St>& = Ghost<dim, St>(0))
        : pos_host_(n * dim, St(0)) {
        for (unsigned d = 0; d < dim; ++d) {
            lo_[d] = box.getLow(d);
            hi_[d] = box.getHigh(d);
            if (bc[d] == PERIODIC) periodic_ |= 1u << d;
        }
        for (auto& p : props_) p.host.assign(n, P(0));
    }

    // ---- construction on the host ---------------------------------------------------
    void add() {
        pos_host_.resize(pos_host_.size() + dim, St(0));
        for (auto& p : props_) p.host.push_back(P(0));
        pos_dirty_ = true;
        for (auto& p : props_) p.host_dirty = true;
    }
    St* getLastPos() {
        pos_dirty_ = true;
        return &pos_host_[pos_host_.size() - dim];
    }
    template <unsigned int prp>
    P& getLastProp() {
        static_assert(prp < nprop, "property index out of range");
        auto& p = host_prop(prp);
        p.host_dirty = true;
        return p.host.back();
    }
    size_t size_local() const { return pos_host_.size() / dim; }

    // one process, one GPU: nothing to redistribute or exchange
    void map() { ++map_ctr_; }
    template <unsigned int... prp>
    void ghost_get() {}
    void ghost_get_subset() {}
    void deleteGhost() {}

    // ---- host access ---------------------------------------------------------------------
    Point<dim, St> getPos(vect_dist_key_dx key) const {
        Point<dim, St> p;
        for (unsigned d = 0; d < dim; ++d) p.get(d) = pos_host_[key.getKey() * dim + d];
        return p;
    }
    St* getPosWrite(size_t key) {
        pos_dirty_ = true;
        ++map_ctr_;
        return &pos_host_[key * dim];
    }
    template <unsigned int prp>
    P& getProp(vect_dist_key_dx key) {
        auto& p = host_prop(prp);
        p.host_dirty = true;
        return p.host[key.getKey()];
    }
    template <unsigned int prp>
    P& getProp(size_t key) {
        return getProp<prp>(vect_dist_key_dx(key));
    }

    struct DomainIterator {
        size_t i, n;
        bool isNext() const { return i < n; }
        vect_dist_key_dx get() const { return vect_dist_key_dx(i); }
        DomainIterator& operator++() {
            ++i;
            return *this;
        }
    };
    DomainIterator getDomainIterator() const { return {0, size_local()}; }
    grid_key_iterator<dim> getGridIterator(const size_t (&sz)[dim]) const { return grid_key_iterator<dim>(sz); }

    // ---- device access (used by the operators) ---------------------------------------------
    const St* devPos() {
        if (pos_dirty_) {
            const size_t n = size_local();
            std::vector<St> soa(n * dim);
            for (size_t i = 0; i < n; ++i)
                for (unsigned d = 0; d < dim; ++d) soa[d * n + i] = pos_host_[i * dim + d];
            pos_dev_.upload(soa);
            pos_dirty_ = false;
        }
        return pos_dev_.data();
    }
    // Device pointer of a property, valid until the next host write.
    P* devProp(unsigned int prp) {
        auto& p = props_.at(prp);
        if (p.host_dirty) {
            p.dev.upload(p.host);
            p.host_dirty = false;
        }
        if (p.dev.size() != size_local()) p.dev.resize(size_local());
        return p.dev.data();
    }
    // The device copy of prp was written; the host copy is refreshed on access.
    void devPropWritten(unsigned int prp) {
        props_.at(prp).dev_dirty = true;
        props_.at(prp).host_dirty = false;
    }
    void hostToDevicePos() { devPos(); }
    template <unsigned int prp>
    void deviceToHostProp() { host_prop(prp); }

    const St* boxLow() const { return lo_; }
    const St* boxHigh() const { return hi_; }
    unsigned periodicMask() const { return periodic_; }
    size_t getMapCtr() const { return map_ctr_; }

    template <unsigned int opt = VL_NON_SYMMETRIC | VL_SKIP_REF_PART>
    VerletList_hip<vector_dist_hip> getVerlet(St rCut) {
        static_assert(opt & VL_SKIP_REF_PART, "DC-PSE supports exclude the particle itself");
        return VerletList_hip<vector_dist_hip>(*this, rCut);
    }
    void updateVerlet(VerletList_hip<vector_dist_hip>& vl, St rCut) { vl = VerletList_hip<vector_dist_hip>(*this, rCut); }

    // Moves particle key to x (dim coordinates); the positions changed, so
    // operators assembled before must be updated (map counter).
    void setPos(size_t key, const St* x) {
        St* p = getPosWrite(key);
        for (unsigned d = 0; d < dim; ++d) p[d] = x[d];
    }

   private:
    struct Prop {
        std::vector<P> host;
        device_array<P> dev;
        bool host_dirty = true, dev_dirty = false;
    };
    Prop& host_prop(unsigned int prp) {
        auto& p = props_.at(prp);
        if (p.dev_dirty) {
            p.host.resize(p.dev.size());
            p.dev.download(p.host.data());
            p.dev_dirty = false;
        }
        return p;
    }
    std::vector<St> pos_host_;
    device_array<St> pos_dev_;
    bool pos_dirty_ = true;
    std::array<Prop, nprop> props_;
    St lo_[dim], hi_[dim];
    unsigned periodic_ = 0;
    size_t map_ctr_ = 0;
};

// The particles ids of a vector_dist_hip, the counterpart of openfpm's
// vector_dist_subset (the domain of two-set operators whose support is the
// whole set, src/DCPSE/DCPSE_op/tests/DCPSE_op_subset_test.cpp:113-141). It
// is a view: positions and properties are the parent's at ids; its Verlet
// list holds, for subset particle p, every parent particle q != ids[p]
// within rCut. The reference selects the subset by a per-particle subset
// number; here the caller gives the indices.
template <typename parent_type>
class vector_dist_subset_hip {
   public:
    static constexpr unsigned int dims = parent_type::dims;
    typedef typename parent_type::stype stype;
    typedef typename parent_type::prop_type prop_type;
    typedef parent_type parent_t;

    vector_dist_subset_hip(parent_type& parent, std::vector<int32_t> ids) : parent_(parent), ids_(std::move(ids)) {
        for (int32_t i : ids_)
            if (i < 0 || (size_t)i >= parent_.size_local()) throw dcpse_error("subset index out of range");
    }
    size_t size_local() const { return ids_.size(); }
    const std::vector<int32_t>& getIds() const { return ids_; }
    // index in the parent set of subset particle p
    size_t supportIndex(size_t p) const { return (size_t)ids_[p]; }
    parent_type& parent() { return parent_; }

    Point<dims, stype> getPos(vect_dist_key_dx key) const { return parent_.getPos(vect_dist_key_dx(ids_[key.getKey()])); }
    template <unsigned int prp>
    prop_type& getProp(vect_dist_key_dx key) {
        return parent_.template getProp<prp>(vect_dist_key_dx((size_t)ids_[key.getKey()]));
    }
    template <unsigned int prp>
    prop_type& getProp(size_t key) {
        return getProp<prp>(vect_dist_key_dx(key));
    }
    typename parent_type::DomainIterator getDomainIterator() const { return {0, size_local()}; }

    // device copies: the subset positions (SoA, gathered again after the
    // parent moved) and the ids
    const stype* devPos() {
        if (pos_ctr_ != parent_.getMapCtr() || pos_dev_.size() != dims * ids_.size()) {
            const size_t n = ids_.size();
            std::vector<stype> soa(n * dims);
            for (size_t i = 0; i < n; ++i) {
                const auto x = parent_.getPos(vect_dist_key_dx((size_t)ids_[i]));
                for (unsigned d = 0; d < dims; ++d) soa[d * n + i] = x[d];
            }
            pos_dev_.upload(soa);
            pos_ctr_ = parent_.getMapCtr();
        }
        return pos_dev_.data();
    }
    const int32_t* devIds() {
        if (ids_dev_.size() != ids_.size()) ids_dev_.upload(ids_);
        return ids_dev_.data();
    }
    const stype* boxLow() const { return parent_.boxLow(); }
    const stype* boxHigh() const { return parent_.boxHigh(); }
    unsigned periodicMask() const { return parent_.periodicMask(); }
    size_t getMapCtr() const { return parent_.getMapCtr(); }

    template <unsigned int opt = VL_NON_SYMMETRIC | VL_SKIP_REF_PART>
    VerletList_hip<parent_type> getVerlet(stype rCut) {
        static_assert(opt & VL_SKIP_REF_PART, "DC-PSE supports exclude the particle itself");
        return VerletList_hip<parent_type>(parent_, *this, rCut);
    }
    void updateVerlet(VerletList_hip<parent_type>& vl, stype rCut) { vl = VerletList_hip<parent_type>(parent_, *this, rCut); }

   private:
    parent_type& parent_;
    std::vector<int32_t> ids_;
    device_array<stype> pos_dev_;
    device_array<int32_t> ids_dev_;
    size_t pos_ctr_ = ~size_t(0);
};

// Neighbour CSR in device memory; row p holds every q != p with |x_p - x_q| < rCut
// in ascending index (dcpse_support_count / dcpse_support_fill), or, for a
// subset domain, every parent particle q != ids[p] (dcpse_support_*_subset).
template <typename vector_type>
class VerletList_hip {
   public:
    typedef typename vector_type::stype St;
    template <typename subset_type>
    VerletList_hip(vector_type& parts, subset_type& dom, St rCut) : n_rows_(dom.size_local()) {
        Context& c = context();
        device_array<double> rc(1);
        double r = (double)rCut;
        rc.upload(&r, 1);
        std::vector<double> lo(vector_type::dims), hi(vector_type::dims);
        for (unsigned d = 0; d < vector_type::dims; ++d) {
            lo[d] = (double)parts.boxLow()[d];
            hi[d] = (double)parts.boxHigh()[d];
        }
        const double* pos = pos_double(parts);
        const double* pdom = dom.devPos();
        const int32_t* ids = dom.devIds();
        row_ptr_.resize(n_rows_ + 1);
        int64_t nnz = 0;
        c.check(dcpse_support_count_subset(c.get(), vector_type::dims, pos, (int64_t)parts.size_local(), pdom, ids,
                                           (int64_t)n_rows_, rc.data(), 0, lo.data(), hi.data(), parts.periodicMask(),
                                           row_ptr_.data(), &nnz),
                "dcpse_support_count_subset");
        col_.resize((size_t)nnz);
        if (nnz)
            c.check(dcpse_support_fill_subset(c.get(), vector_type::dims, pos, (int64_t)parts.size_local(), pdom, ids,
                                              (int64_t)n_rows_, rc.data(), 0, lo.data(), hi.data(),
                                              parts.periodicMask(), row_ptr_.data(), col_.data()),
                    "dcpse_support_fill_subset");
    }
    VerletList_hip(vector_type& parts, St rCut) : n_rows_(parts.size_local()) {
        Context& c = context();
        device_array<double> rc(1);
        double r = (double)rCut;
        rc.upload(&r, 1);
        std::vector<double> lo(vector_type::dims), hi(vector_type::dims);
        for (unsigned d = 0; d < vector_type::dims; ++d) {
            lo[d] = (double)parts.boxLow()[d];
            hi[d] = (double)parts.boxHigh()[d];
        }
        const double* pos = pos_double(parts);
        row_ptr_.resize(n_rows_ + 1);
        int64_t nnz = 0;
        c.check(dcpse_support_count(c.get(), vector_type::dims, pos, (int64_t)n_rows_, nullptr, (int64_t)n_rows_,
                                    rc.data(), 0, lo.data(), hi.data(), parts.periodicMask(), row_ptr_.data(), &nnz),
                "dcpse_support_count");
        col_.resize((size_t)nnz);
        if (nnz)
            c.check(dcpse_support_fill(c.get(), vector_type::dims, pos, (int64_t)n_rows_, nullptr, (int64_t)n_rows_,
                                       rc.data(), 0, lo.data(), hi.data(), parts.periodicMask(), row_ptr_.data(),
                                       col_.data()),
                    "dcpse_support_fill");
    }
    size_t getNNPart(vect_dist_key_dx p) { return mirror_rp()[p.getKey() + 1] - mirror_rp()[p.getKey()]; }
    size_t getNNPart(size_t p) { return getNNPart(vect_dist_key_dx(p)); }
    size_t get(vect_dist_key_dx p, size_t j) {
        mirror();
        return (size_t)col_host_[rp_host_[p.getKey()] + j];
    }
    size_t get(size_t p, size_t j) { return get(vect_dist_key_dx(p), j); }
    size_t size() const { return n_rows_; }
    size_t nnz() const { return col_.size(); }
    const int64_t* devRowPtr() const { return row_ptr_.data(); }
    const int32_t* devCol() const { return col_.data(); }
    // host copies of the CSR (read back once)
    const std::vector<int64_t>& hostRowPtr() { return mirror_rp(); }
    const std::vector<int32_t>& hostCol() {
        mirror();
        return col_host_;
    }

    // The sliced-ELL structure of this support (dcpse_sell_count: slices of 64
    // rows, lane -> row) and, with cols, its 32-bit column array
    // (dcpse_sell_fill_cols); built on first use and shared by every operator
    // on this list, which keep only their own weights in its slot order.
    // Throws dcpse_error when device memory runs out.
    struct sell_layout {
        device_array<int64_t> slice_ptr;
        device_array<int32_t> sell_row, sell_col;
        int64_t n_slots = -1;
    };
    const sell_layout& sell(bool cols) {
        Context& c = context();
        const int64_t n = (int64_t)n_rows_;
        if (sell_.n_slots < 0) {
            const int64_t n_lanes = 64 * ((n + 63) / 64);
            sell_layout L;
            L.slice_ptr.resize((size_t)(n_lanes / 64 + 1));
            L.sell_row.resize((size_t)std::max<int64_t>(n_lanes, 64));
            int64_t n_slots = 0;
            c.check(dcpse_sell_count(c.get(), n, row_ptr_.data(), L.slice_ptr.data(), L.sell_row.data(), &n_slots),
                    "dcpse_sell_count");
            L.n_slots = n_slots;
            sell_ = std::move(L);
        }
        if (cols && sell_.sell_col.size() == 0) {
            device_array<int32_t> sc((size_t)std::max<int64_t>(sell_.n_slots, 4));
            if (n)
                c.check(dcpse_sell_fill_cols(c.get(), n, row_ptr_.data(), col_.data(), sell_.slice_ptr.data(),
                                             sell_.sell_row.data(), sc.data()),
                        "dcpse_sell_fill_cols");
            sell_.sell_col = std::move(sc);
        }
        return sell_;
    }

   private:
    static const double* pos_double(vector_type& parts) {
        static_assert(std::is_same<St, double>::value, "positions are double precision");
        return parts.devPos();
    }
    const std::vector<int64_t>& mirror_rp() {
        mirror();
        return rp_host_;
    }
    void mirror() {
        if (rp_host_.empty()) {
            rp_host_ = row_ptr_.to_host();
            col_host_ = col_.to_host();
        }
    }
    size_t n_rows_;
    device_array<int64_t> row_ptr_;
    device_array<int32_t> col_;
    std::vector<int64_t> rp_host_;
    std::vector<int32_t> col_host_;
    sell_layout sell_;
};

}  // namespace openfpm_dcpse
