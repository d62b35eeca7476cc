/*
 * oops_classes.hpp — the reference's C++ class interfaces of the expv path,
 * header-only, over the C ABI of oops_expv.h / oops_cce.h.
 *
 * A caller of the reference's classes includes this header instead of
 * include/math/MatExp.h, include/kron/KronProd.h, include/spin/SpinCluster.h
 * and include/spin/SpinClusterAlgorithm.h and keeps its code: the class,
 * method and enum names, argument meanings and the column-major layout of
 * every matrix are the reference's.  Armadillo is replaced by the small
 * column-major containers below (Mat, Col: the memory layout of arma::Mat /
 * arma::Col), so memptr() hands the same bytes to the C ABI.  Errors throw
 * std::runtime_error carrying oops_last_error() where the reference stops the
 * process.
 *
 * Reference interfaces mirrored (paths relative to the reference's src/):
 *   KronProd, SumKronProd, Flat/Sharp/CircleC, Expand  include/kron/KronProd.h:17-104,
 *                                                      source/kron/KronProd.cpp:9-231
 *   MatExpVector                                       include/math/MatExp.h:45-87,
 *                                                      source/math/MatExp.cpp:86-521
 *   cClusterIndex                                      include/spin/SpinClusterIndex.h:15-40
 *   cSpinCollection (coordinates and connection matrix) include/spin/SpinCollection.h,
 *                                                      source/spin/SpinCollection/SpinCollection.cpp:31-57
 *   cSpinGrouping, cDepthFirstPathTracing              include/spin/SpinClusterAlgorithm.h:36-82
 *   cSpinCluster                                       include/spin/SpinCluster.h:31-68
 */
#ifndef OOPS_CLASSES_HPP
#define OOPS_CLASSES_HPP

#include <algorithm>
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "oops_cce.h"
#include "oops_expv.h"

namespace oops {

typedef std::complex<double> cx_double;

// ---------------------------------------------------------------------------
// Column-major dense matrix (the element layout of arma::Mat<T>).
template <typename T>
class Mat {
public:
    size_t n_rows = 0, n_cols = 0, n_elem = 0;

    Mat() = default;
    Mat(size_t r, size_t c) : n_rows(r), n_cols(c), n_elem(r * c), mem_(r * c, T(0)) {}
    Mat(const T* mem, size_t r, size_t c) : n_rows(r), n_cols(c), n_elem(r * c), mem_(mem, mem + r * c) {}

    static Mat eye(size_t n) {
        Mat I(n, n);
        for (size_t i = 0; i < n; ++i) I(i, i) = T(1);
        return I;
    }
    T& operator()(size_t i, size_t j) { return mem_[i + n_rows * j]; }
    const T& operator()(size_t i, size_t j) const { return mem_[i + n_rows * j]; }
    T& operator()(size_t k) { return mem_[k]; }
    const T& operator()(size_t k) const { return mem_[k]; }
    T& operator[](size_t k) { return mem_[k]; }
    const T& operator[](size_t k) const { return mem_[k]; }
    T* memptr() { return mem_.data(); }
    const T* memptr() const { return mem_.data(); }

    // transpose without conjugation (arma .st())
    Mat st() const {
        Mat t(n_cols, n_rows);
        for (size_t j = 0; j < n_cols; ++j)
            for (size_t i = 0; i < n_rows; ++i) t(j, i) = (*this)(i, j);
        return t;
    }
    Mat& operator+=(const Mat& o) {
        check_same(o);
        for (size_t k = 0; k < n_elem; ++k) mem_[k] += o.mem_[k];
        return *this;
    }
    Mat& operator-=(const Mat& o) {
        check_same(o);
        for (size_t k = 0; k < n_elem; ++k) mem_[k] -= o.mem_[k];
        return *this;
    }
    Mat& operator*=(T s) {
        for (auto& x : mem_) x *= s;
        return *this;
    }

private:
    void check_same(const Mat& o) const {
        if (o.n_rows != n_rows || o.n_cols != n_cols) throw std::invalid_argument("matrix sizes differ");
    }
    std::vector<T> mem_;
};

template <typename T>
Mat<T> operator+(Mat<T> a, const Mat<T>& b) { return a += b; }
template <typename T>
Mat<T> operator-(Mat<T> a, const Mat<T>& b) { return a -= b; }
template <typename T>
Mat<T> operator*(T s, Mat<T> a) { return a *= s; }

// Column vector: an n x 1 matrix.
template <typename T>
class Col : public Mat<T> {
public:
    Col() = default;
    explicit Col(size_t n) : Mat<T>(n, 1) {}
    Col(const T* mem, size_t n) : Mat<T>(mem, n, 1) {}
    Col(std::initializer_list<T> v) : Mat<T>(v.begin(), v.size(), 1) {}
    Col(const Mat<T>& m) : Mat<T>(m) {
        if (m.n_cols != 1) throw std::invalid_argument("not a column");
    }
};

typedef Mat<cx_double> cx_mat;
typedef Mat<double> mat;
typedef Col<cx_double> cx_vec;
typedef Col<double> vec;
typedef Mat<size_t> umat;
typedef Col<size_t> uvec;

inline cx_mat conj(const cx_mat& m) {
    cx_mat c(m);
    for (size_t k = 0; k < c.n_elem; ++k) c(k) = std::conj(c(k));
    return c;
}

// Kronecker product, column-major: (A (x) B)(ra nb + rb, ca nb + cb) = A(ra, ca) B(rb, cb).
inline cx_mat kron(const cx_mat& A, const cx_mat& B) {
    cx_mat C(A.n_rows * B.n_rows, A.n_cols * B.n_cols);
    for (size_t ca = 0; ca < A.n_cols; ++ca)
        for (size_t ra = 0; ra < A.n_rows; ++ra)
            for (size_t cb = 0; cb < B.n_cols; ++cb)
                for (size_t rb = 0; rb < B.n_rows; ++rb)
                    C(ra * B.n_rows + rb, ca * B.n_cols + cb) = A(ra, ca) * B(rb, cb);
    return C;
}

// Sparse complex matrix in coordinate form (the entries of arma::sp_cx_mat).
struct sp_cx_mat {
    size_t n_rows = 0, n_cols = 0;
    std::vector<size_t> row, col;   // 0-based
    std::vector<cx_double> val;
    sp_cx_mat() = default;
    sp_cx_mat(size_t r, size_t c) : n_rows(r), n_cols(c) {}
    void insert(size_t i, size_t j, cx_double v) {
        row.push_back(i);
        col.push_back(j);
        val.push_back(v);
    }
};

// ---------------------------------------------------------------------------
// Kronecker products (KronProd.h:17-104; KronProd.cpp:9-231).
typedef std::vector<size_t> INDICES;
typedef std::vector<cx_mat> TERM;
typedef double MULTIPLIER;
typedef std::vector<size_t> DIM_LIST;

inline cx_mat Flat(const cx_mat& m) { return kron(cx_mat::eye(m.n_rows), m); }
inline cx_mat Sharp(const cx_mat& m) { return kron(m.st(), cx_mat::eye(m.n_rows)); }
inline cx_mat CircleC(const cx_mat& m) { return kron(cx_mat::eye(m.n_rows), m) - kron(conj(m), cx_mat::eye(m.n_rows)); }
typedef cx_mat(MatExpanFunc)(const cx_mat&);
inline MatExpanFunc* const FLAT = &Flat;
inline MatExpanFunc* const SHARP = &Sharp;
inline MatExpanFunc* const CIRCLEC = &CircleC;

class KronProd {
public:
    KronProd() = default;
    explicit KronProd(DIM_LIST dim_list) : _kron_num(dim_list.size()), _dim_list(dim_list), _dim(1) {
        for (size_t d : _dim_list) _dim *= d;
    }

    // coeff * (x)_s M_s with M_s the term's matrix at its spins, the identity elsewhere
    cx_mat full() const {
        cx_mat r = cx_mat::eye(1);
        for (size_t s = 0; s < _dim_list.size(); ++s) r = kron(r, local(s));
        return cx_double(_coeff, 0.0) * r;
    }
    // coeff * (x)_s vec(M_s)
    cx_vec vecterize() const {
        cx_mat r = cx_mat::eye(1);
        for (size_t s = 0; s < _dim_list.size(); ++s) {
            const cx_mat m = local(s);
            r = kron(r, cx_mat(m.memptr(), m.n_elem, 1));
        }
        return cx_vec(cx_double(_coeff, 0.0) * r);
    }
    void fill(INDICES idx, MULTIPLIER coeff, TERM mat) {
        if (idx.size() != mat.size()) throw std::invalid_argument("KronProd::fill: one matrix per index");
        for (size_t b = 0; b < idx.size(); ++b)
            if (idx[b] >= _dim_list.size() || mat[b].n_rows != _dim_list[idx[b]] || mat[b].n_cols != mat[b].n_rows)
                throw std::invalid_argument("KronProd::fill: matrix does not match its spin's dimension");
        _spin_index = idx;
        _coeff = coeff;
        _mat = mat;
    }
    KronProd& scale(double factor) {
        _coeff *= factor;
        return *this;
    }
    DIM_LIST getDimList() const { return _dim_list; }
    size_t getDim() const { return _dim; }
    MULTIPLIER getCoeff() const { return _coeff; }
    INDICES getIndices() const { return _spin_index; }
    TERM getTermMat() const { return _mat; }
    size_t getKronNum() const { return _kron_num; }
    size_t getMatNum() const { return _mat.size(); }

    friend KronProd Expand(const KronProd& kp, MatExpanFunc* f) {
        DIM_LIST nd;
        for (size_t d : kp._dim_list) nd.push_back(d * d);
        KronProd r(nd);
        TERM nm;
        for (const cx_mat& A : kp._mat) nm.push_back(f(A));
        r.fill(kp._spin_index, kp._coeff, nm);
        return r;
    }

private:
    cx_mat local(size_t s) const {
        for (size_t b = 0; b < _spin_index.size(); ++b)
            if (_spin_index[b] == s) return _mat[b];
        return cx_mat::eye(_dim_list[s]);
    }
    size_t _kron_num = 0;
    DIM_LIST _dim_list;
    INDICES _spin_index;
    MULTIPLIER _coeff = 0.0;
    TERM _mat;
    size_t _dim = 1;
};

class SumKronProd {
public:
    SumKronProd() = default;
    explicit SumKronProd(const std::vector<KronProd>& kp_lst) : _kron_prod_list(kp_lst) {
        if (!kp_lst.empty()) {
            _dim_list = kp_lst[0].getDimList();
            _dim = kp_lst[0].getDim();
            _kron_num = kp_lst[0].getKronNum();
        }
    }
    cx_mat full() const {
        cx_mat r = _kron_prod_list.at(0).full();
        for (size_t i = 1; i < _kron_prod_list.size(); ++i) r += _kron_prod_list[i].full();
        return r;
    }
    cx_vec vecterize() const {
        cx_mat r = _kron_prod_list.at(0).vecterize();
        for (size_t i = 1; i < _kron_prod_list.size(); ++i) r += _kron_prod_list[i].vecterize();
        return cx_vec(r);
    }
    std::vector<KronProd> getKronProdList() const { return _kron_prod_list; }
    DIM_LIST getDimList() const { return _dim_list; }
    size_t getDim() const { return _dim; }
    size_t getKronProdSize() const { return _kron_prod_list.size(); }
    size_t getKronNum() const { return _kron_num; }
    SumKronProd& scale(double factor) {
        for (auto& kp : _kron_prod_list) kp.scale(factor);
        return *this;
    }
    void append(KronProd kp) { _kron_prod_list.push_back(kp); }
    std::vector<MULTIPLIER> getCoeffList() const {
        std::vector<MULTIPLIER> r;
        for (const auto& kp : _kron_prod_list) r.push_back(kp.getCoeff());
        return r;
    }
    std::vector<INDICES> getIndicesList() const {
        std::vector<INDICES> r;
        for (const auto& kp : _kron_prod_list) r.push_back(kp.getIndices());
        return r;
    }
    std::vector<TERM> getTermList() const {
        std::vector<TERM> r;
        for (const auto& kp : _kron_prod_list) r.push_back(kp.getTermMat());
        return r;
    }
    std::vector<size_t> getMatNumList() const {
        std::vector<size_t> r;
        for (const auto& kp : _kron_prod_list) r.push_back(kp.getMatNum());
        return r;
    }
    friend SumKronProd Expand(const SumKronProd& skp, MatExpanFunc* f) {
        std::vector<KronProd> l;
        for (const auto& kp : skp._kron_prod_list) l.push_back(Expand(kp, f));
        return SumKronProd(l);
    }
    friend SumKronProd& operator+(SumKronProd& sum, const SumKronProd skp) {
        sum._kron_prod_list.insert(sum._kron_prod_list.end(), skp._kron_prod_list.begin(),
                                   skp._kron_prod_list.end());
        if (sum._dim_list.empty()) {
            sum._dim_list = skp._dim_list;
            sum._dim = skp._dim;
            sum._kron_num = skp._kron_num;
        }
        return sum;
    }

protected:
    size_t _dim = 0;
    size_t _kron_num = 0;
    DIM_LIST _dim_list;
    std::vector<KronProd> _kron_prod_list;
};

// ---------------------------------------------------------------------------
// One HIP context per process for the class layer (device 0, as the
// reference's expv library runs on the current device).
inline oops_ctx* class_context() {
    static oops_ctx* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (oops_ctx_create(0, &ctx) != OOPS_OK) ctx = nullptr;
    });
    if (!ctx) throw std::runtime_error("oops: no HIP device");
    return ctx;
}

inline void check_status(int rc, const char* what) {
    if (rc != OOPS_OK) {
        std::string msg = std::string(what) + " failed with status " + std::to_string(rc);
        oops_ctx* c = nullptr;
        try { c = class_context(); } catch (...) {}
        if (c) msg += std::string(": ") + oops_last_error(c);
        throw std::runtime_error(msg);
    }
}

// ---------------------------------------------------------------------------
// MatExpVector (MatExp.h:45-87): w(t_k) = exp(prefactor A t_k) v for every t_k.
class MatExpVector {
public:
    enum MatExpVectorMethod { Explicit, ExplicitSparse, Inexplicit, InexplicitGPU };

    MatExpVector() = default;
    // H as a sum of Kronecker products; the prefactor is -i (MatExp.cpp:86-103).
    // Inexplicit and InexplicitGPU flatten it into the reference's term arrays
    // and run the device Krylov expv (oops_expv); Explicit and ExplicitSparse
    // expand it into a dense / sparse matrix first, as the reference does.
    MatExpVector(const SumKronProd& skp, const cx_vec& v, const vec& time_list, MatExpVectorMethod method)
        : _method(method), _skp(skp), _vector(v), _prefactor(0.0, -1.0), _time_list(time_list),
          _nTime(time_list.n_elem), _dim(skp.getDim()) {}
    // dense A: the generator is prefactor * A (MatExp.cpp:104-120)
    MatExpVector(const cx_mat& m, const cx_vec& v, const cx_double prefactor, const vec& time_list)
        : _method(Explicit), _matrix(prefactor * m), _vector(v), _prefactor(prefactor), _time_list(time_list),
          _nTime(time_list.n_elem), _dim(m.n_cols) {}
    // sparse A (MatExp.cpp:121-137)
    MatExpVector(const sp_cx_mat& m, const cx_vec& v, const cx_double prefactor, const vec& time_list)
        : _method(ExplicitSparse), _sp_matrix(m), _vector(v), _prefactor(prefactor), _time_list(time_list),
          _nTime(time_list.n_elem), _dim(m.n_cols) {
        for (auto& z : _sp_matrix.val) z *= prefactor;
    }

    cx_mat run() {
        switch (_method) {
            case Explicit: return runExplicit();
            case ExplicitSparse: return runExplicitSparse();
            case Inexplicit: return runInexplicit();
            case InexplicitGPU: return runInexplicitGPU();
        }
        throw std::invalid_argument("MatExpVector: unknown method");
    }

    // krylov_zgexpv on the dense generator (MatExp.cpp:162-178)
    cx_mat runExplicit() {
        if (_matrix.n_elem == 0 && _skp.getKronProdSize() > 0) _matrix = _prefactor * _skp.full();
        check_vector();
        _resVectorList = cx_mat(_dim, _nTime);
        const int rc = krylov_zgexpv((int)_dim, reinterpret_cast<const double*>(_matrix.memptr()),
                                     reinterpret_cast<const double*>(_vector.memptr()), (int)_nTime,
                                     _time_list.memptr(), reinterpret_cast<double*>(_resVectorList.memptr()),
                                     (int)_klim, (int)_krylov_m, _krylov_tol, (int)_itrace);
        check_status(rc, "krylov_zgexpv");
        return _resVectorList;
    }
    // krylov_zcooexpv on the sparse generator, 1-based coordinates (MatExp.cpp:180-214)
    cx_mat runExplicitSparse() {
        if (_sp_matrix.val.empty() && _skp.getKronProdSize() > 0) {
            const cx_mat f = _prefactor * _skp.full();
            _sp_matrix = sp_cx_mat(f.n_rows, f.n_cols);
            for (size_t j = 0; j < f.n_cols; ++j)
                for (size_t i = 0; i < f.n_rows; ++i)
                    if (f(i, j) != cx_double(0.0)) _sp_matrix.insert(i, j, f(i, j));
        }
        check_vector();
        std::vector<int> ia, ja;
        for (size_t e = 0; e < _sp_matrix.val.size(); ++e) {
            ia.push_back((int)_sp_matrix.row[e] + 1);
            ja.push_back((int)_sp_matrix.col[e] + 1);
        }
        _resVectorList = cx_mat(_dim, _nTime);
        const int rc = krylov_zcooexpv((int)_dim, (int)_sp_matrix.val.size(), ia.data(), ja.data(),
                                       reinterpret_cast<const double*>(_sp_matrix.val.data()),
                                       reinterpret_cast<const double*>(_vector.memptr()), (int)_nTime,
                                       _time_list.memptr(), reinterpret_cast<double*>(_resVectorList.memptr()),
                                       (int)_klim, (int)_krylov_m, _krylov_tol, (int)_itrace);
        check_status(rc, "krylov_zcooexpv");
        return _resVectorList;
    }
    // The reference's CPU Krylov driver (MatExp.cpp:216-365); here the device runs it.
    cx_mat runInexplicit() { return runInexplicitGPU(); }
    // Flattened term arrays -> device Krylov expv (MatExp.cpp:367-521 / MAIN_CACHE)
    cx_mat runInexplicitGPU() {
        check_vector();
        Flattened f(_skp);
        _resVectorList = cx_mat(_dim, _nTime);
        const int rc = oops_expv(class_context(), &f.table, reinterpret_cast<const double*>(_vector.memptr()),
                                 (int32_t)_nTime, _time_list.memptr(), (int32_t)_klim, (int32_t)_krylov_m, _krylov_tol,
                                 0, reinterpret_cast<double*>(_resVectorList.memptr()), nullptr);
        check_status(rc, "oops_expv");
        return _resVectorList;
    }
    cx_mat getResult() const { return _resVectorList; }

    void enable_step_print() { _itrace = 1; }
    void enable_skp_print() { _is_print_skp = true; }
    void disable_step_print() { _itrace = 0; }
    void disable_skp_print() { _is_print_skp = false; }

private:
    // The flattened arrays MatExpVector builds (MatExp.cpp:219-262).
    struct Flattened {
        std::vector<double> coeff;
        std::vector<int64_t> nbody, pos_off, pos, dims, mat_off, spin_dim;
        std::vector<cx_double> mats;
        oops_term_table table;
        explicit Flattened(const SumKronProd& skp) {
            pos_off.push_back(0);
            mat_off.push_back(0);
            for (const KronProd& kp : skp.getKronProdList()) {
                coeff.push_back(kp.getCoeff());
                const INDICES idx = kp.getIndices();
                const TERM t = kp.getTermMat();
                nbody.push_back((int64_t)idx.size());
                for (size_t b = 0; b < idx.size(); ++b) {
                    pos.push_back((int64_t)idx[b]);
                    dims.push_back((int64_t)t[b].n_rows);
                    mat_off.push_back(mat_off.back() + (int64_t)t[b].n_elem);
                    mats.insert(mats.end(), t[b].memptr(), t[b].memptr() + t[b].n_elem);
                }
                pos_off.push_back((int64_t)pos.size());
            }
            for (size_t d : skp.getDimList()) spin_dim.push_back((int64_t)d);
            if (pos.empty()) { pos.push_back(0); dims.push_back(1); }
            if (mats.empty()) mats.push_back(0.0);
            table.nspin = (int32_t)spin_dim.size();
            table.nterm = (int32_t)coeff.size();
            table.coeff = coeff.data();
            table.nbody = nbody.data();
            table.pos_offset = pos_off.data();
            table.pos = pos.data();
            table.dims = dims.data();
            table.mat_offset = mat_off.data();
            table.mats = reinterpret_cast<const double*>(mats.data());
            table.spin_dim = spin_dim.data();
        }
    };
    void check_vector() const {
        if (_vector.n_elem != _dim) throw std::invalid_argument("MatExpVector: vector length differs from the dimension");
    }

    MatExpVectorMethod _method = Inexplicit;
    SumKronProd _skp;
    cx_mat _matrix;
    sp_cx_mat _sp_matrix;
    cx_vec _vector;
    cx_double _prefactor{0.0, -1.0};
    vec _time_list;
    size_t _nTime = 0;
    size_t _dim = 0;
    cx_mat _resVectorList;
    // MatExp.cpp:98-101
    size_t _klim = 10;
    size_t _krylov_m = 30;
    double _krylov_tol = 1e-12;
    size_t _itrace = 0;
    bool _is_print_skp = false;
};

// ---------------------------------------------------------------------------
// Clusters (SpinClusterIndex.h, SpinCollection.h, SpinClusterAlgorithm.h, SpinCluster.h).
typedef std::pair<size_t, size_t> ClusterPostion;

class cClusterIndex {
public:
    cClusterIndex() = default;
    explicit cClusterIndex(const uvec& idx, std::vector<size_t> sub = {})
        : _index(idx), _spin_num(idx.n_elem), _sub_clst_pos(std::move(sub)) {}
    uvec getIndex() const { return _index; }
    size_t getNum() const { return _spin_num; }
    size_t getOrder() const { return _spin_num - 1; }
    // positions (in the order below) of the parent clusters that generated it
    std::set<ClusterPostion> getSubClstPos() const {
        std::set<ClusterPostion> r;
        for (size_t p : _sub_clst_pos) r.insert({getOrder() - 1, p});
        return r;
    }

private:
    uvec _index;
    size_t _spin_num = 0;
    std::vector<size_t> _sub_clst_pos;
};

// Spin coordinates (Angstrom) and the pair graph d(i, j) <= threshold.
class cSpinCollection {
public:
    cSpinCollection() = default;
    explicit cSpinCollection(const std::vector<std::array<double, 3>>& coords) : _coords(coords) {}
    size_t getSpinNum() const { return _coords.size(); }
    const std::vector<std::array<double, 3>>& getCoordinates() const { return _coords; }
    // dense 0/1 connection matrix (SpinCollection.cpp:38-55: d computed for
    // j >= i and mirrored, diagonal zeroed)
    mat getConnectionMatrix(double threshold) const {
        const size_t n = _coords.size();
        mat c(n, n);
        for (size_t i = 0; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) {
                const double dx = _coords[i][0] - _coords[j][0], dy = _coords[i][1] - _coords[j][1],
                             dz = _coords[i][2] - _coords[j][2];
                double a = dx * dx, b = dy * dy;
                a += dz * dz;
                if (std::sqrt(a + b) <= threshold) c(i, j) = c(j, i) = 1.0;
            }
        return c;
    }

private:
    std::vector<std::array<double, 3>> _coords;
};

class cSpinGrouping {
public:
    cSpinGrouping() = default;
    cSpinGrouping(const mat& connection_matrix, size_t max_order)
        : _nspin(connection_matrix.n_rows), _max_order(max_order), _connection_matrix(connection_matrix) {}
    virtual ~cSpinGrouping() {
        if (_clusters) oops_clusters_destroy(_clusters);
    }
    cSpinGrouping(const cSpinGrouping&) = delete;
    cSpinGrouping& operator=(const cSpinGrouping&) = delete;
    virtual void generate() = 0;
    size_t getMaxOrder() const { return _max_order; }
    const oops_clusters* handle() const { return _clusters; }

protected:
    size_t _nspin = 0;
    size_t _max_order = 0;
    mat _connection_matrix;
    oops_clusters* _clusters = nullptr;
};

// Connected induced subgraphs of 1..maxOrder vertices in the reference's set
// order (size, then lexicographic) with their parent positions
// (SpinClusterAlgorithm.cpp:98-148), enumerated by the library.
class cDepthFirstPathTracing : public cSpinGrouping {
public:
    cDepthFirstPathTracing() = default;
    cDepthFirstPathTracing(const mat& connection_matrix, size_t maxOrder)
        : cSpinGrouping(connection_matrix, maxOrder) {}
    void generate() override {
        if (_clusters) { oops_clusters_destroy(_clusters); _clusters = nullptr; }
        std::vector<int64_t> off(1, 0);
        std::vector<int32_t> nbr;
        for (size_t i = 0; i < _nspin; ++i) {
            for (size_t j = 0; j < _nspin; ++j)
                if (j != i && _connection_matrix(i, j) != 0.0) nbr.push_back((int32_t)j);
            off.push_back((int64_t)nbr.size());
        }
        check_status(oops_clusters_enumerate_graph((int32_t)_nspin, off.data(), nbr.data(), (int32_t)_max_order,
                                                   &_clusters),
                     "oops_clusters_enumerate_graph");
    }
};

class cSpinCluster {
public:
    cSpinCluster() = default;
    cSpinCluster(const cSpinCollection& sc, cSpinGrouping* grouping) : _spin_collection(sc), _grouping(grouping) {}

    // runs the grouping and reads its clusters (cSpinCluster::make)
    void make() {
        if (!_grouping) throw std::invalid_argument("cSpinCluster: no grouping");
        _grouping->generate();
        const oops_clusters* C = _grouping->handle();
        _max_order = _grouping->getMaxOrder();
        _index.assign(_max_order, {});
        _parents.assign(_max_order, {});
        for (size_t o = 0; o < _max_order; ++o) {
            int64_t cnt = 0;
            check_status(oops_clusters_count(C, (int32_t)o, &cnt), "oops_clusters_count");
            std::vector<int32_t> idx((size_t)cnt * (o + 1));
            if (cnt) check_status(oops_clusters_get(C, (int32_t)o, idx.data()), "oops_clusters_get");
            std::vector<int64_t> off((size_t)cnt + 1);
            check_status(oops_clusters_parents(C, (int32_t)o, off.data(), nullptr), "oops_clusters_parents");
            std::vector<int32_t> par((size_t)off[cnt]);
            check_status(oops_clusters_parents(C, (int32_t)o, off.data(), par.data()), "oops_clusters_parents");
            for (int64_t c = 0; c < cnt; ++c) {
                _index[o].emplace_back(idx.begin() + c * (o + 1), idx.begin() + (c + 1) * (o + 1));
                _parents[o].emplace_back(par.begin() + off[c], par.begin() + off[c + 1]);
            }
        }
    }

    size_t getMaxOrder() const { return _max_order; }
    size_t getClusterNum(int order) const { return _index.at(order).size(); }
    // all clusters of one order: count x (order+1) spin indices
    umat getClusterIndex(size_t order) const {
        const auto& l = _index.at(order);
        umat m(l.size(), order + 1);
        for (size_t c = 0; c < l.size(); ++c)
            for (size_t q = 0; q <= order; ++q) m(c, q) = (size_t)l[c][q];
        return m;
    }
    cClusterIndex getClusterIndex(size_t order, size_t index) const {
        const auto& c = _index.at(order).at(index);
        uvec u(c.size());
        for (size_t q = 0; q < c.size(); ++q) u(q) = (size_t)c[q];
        std::vector<size_t> sub(_parents[order][index].begin(), _parents[order][index].end());
        return cClusterIndex(u, sub);
    }
    cClusterIndex getClusterIndex(const ClusterPostion& pos) const { return getClusterIndex(pos.first, pos.second); }
    std::vector<std::array<double, 3>> getClusterCoord(size_t order, size_t index) const {
        std::vector<std::array<double, 3>> r;
        for (int32_t s : _index.at(order).at(index)) r.push_back(_spin_collection.getCoordinates().at(s));
        return r;
    }
    // every connected proper sub-cluster, reached through the parent links
    // (cSpinCluster::getSubClusters, SpinCluster.cpp:82-98)
    std::set<ClusterPostion> getSubClusters(size_t order, size_t index) const {
        std::set<ClusterPostion> r;
        if (order == 0) return r;
        for (int32_t p : _parents.at(order).at(index)) {
            r.insert({order - 1, (size_t)p});
            const auto s = getSubClusters(order - 1, (size_t)p);
            r.insert(s.begin(), s.end());
        }
        return r;
    }

private:
    cSpinCollection _spin_collection;
    cSpinGrouping* _grouping = nullptr;
    size_t _max_order = 0;
    std::vector<std::vector<std::vector<int32_t>>> _index, _parents;
};

}  // namespace oops

#endif  // OOPS_CLASSES_HPP
