// CCE host pipeline (product side, C++17, no Armadillo):
//   * spin operators and isotope constants   (source/spin/Spin/Spin.cpp:28-95, SpinData.cpp:4-17)
//   * dipolar / Zeeman / hyperfine-field couplings (source/misc/misc.cpp:11-57)
//   * term tables in the reference's order and with its exact "!= 0" filter
//     (SpinInteractionComponent.cpp:184-304, SpinInteraction.cpp:29-65)
//   * Liouvillians lvA = H0# - H1b + G, lvB = H1# - H0b + G
//     (source/app/ensemble_cce.cpp:33-91, source/kron/KronProd.cpp:9-28, 84-107)
//   * density vector and observable (SpinState.cpp:9-23, KronProd.cpp:68-82,
//     ensemble_cce.cpp:104-113)
//   * NV electron eigenstates (source/app/DefectCenter.cpp:14-45)
//   * cluster enumeration with the reference's set order and parent lists
//     (SpinCollection.cpp:31-57, SpinClusterAlgorithm.cpp:28-148,
//      SpinClusterIndex.cpp:55-64, SpinCluster.cpp:82-98)
//   * CCE reduction and product (source/app/cce.cpp:151-189), plus a
//     sign/log-magnitude variant that cannot underflow.
//   * a seeded diamond-lattice 13C bath (the synthetic input of the
//     benchmark workloads; not a reference routine).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/oops_cce.h"
#include "../../include/oops_expv.h"
#include "host_common.h"
#include "term_builder.h"

namespace {

typedef std::complex<double> cplx;
typedef std::vector<cplx> Mat;  // column-major square matrix

const double kHbar = 1.0545718e-34;
const double kMu0 = 1.25663706143592e-06;
const double kPi = 3.14159265358979323846;
const double kDistEps = 1e-10;
const double kNVLattice = 3.57;  // DefectCenter.cpp:16

struct Isotope {
    int mult;
    double gamma, omegaQ, eta;
};

bool isotope_of(const std::string& name, Isotope& out) {
    // SpinData.cpp:4-17 (the reference keys the nitrogen data as "14C"/"15C")
    static const std::map<std::string, Isotope> table = {
        {"13C", {2, 6.728284e7, 0.0, 0.0}},
        {"14C", {3, 1.9337792e7, 2.96, 0.0}},
        {"15C", {2, -2.71261804e7, 0.0, 0.0}},
        {"E", {2, -1.760859708e11, 0.0, 0.0}},
        {"NVe", {3, -1.760859708e11, 2.87e3, 0.0}},
    };
    auto it = table.find(name);
    if (it == table.end()) return false;
    out = it->second;
    return true;
}

struct SpinH {
    double r[3];
    Isotope iso;
};

void spin_ops(int mult, Mat& sx, Mat& sy, Mat& sz) {
    const cplx I(0.0, 1.0);
    if (mult == 2) {
        sx = {0.0, 0.5, 0.5, 0.0};
        sy = {0.0, 0.5 * I, -0.5 * I, 0.0};
        sz = {0.5, 0.0, 0.0, -0.5};
    } else {
        const double r = 1.0 / std::sqrt(2.0);
        sx = {0.0, r, 0.0, r, 0.0, r, 0.0, r, 0.0};
        sy = {0.0, r * I, 0.0, -r * I, 0.0, r * I, 0.0, -r * I, 0.0};
        sz = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0};
    }
}

Mat matmul(const Mat& A, const Mat& B, int d) {
    Mat C(d * d, 0.0);
    for (int j = 0; j < d; ++j)
        for (int l = 0; l < d; ++l)
            for (int i = 0; i < d; ++i) C[i + d * j] += A[i + d * l] * B[l + d * j];
    return C;
}

// Euclidean distance summed as ((dx^2 + dz^2) + dy^2), the pairing order of
// Armadillo's direct 2-norm for a 3-vector.
double dist3(const double* a, const double* b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    double acc1 = dx * dx, acc2 = dy * dy;
    acc1 += dz * dz;
    return std::sqrt(acc1 + acc2);
}

std::array<double, 9> dipole(const SpinH& s1, const SpinH& s2) {
    std::array<double, 9> res{};
    const double d = dist3(s1.r, s2.r);
    if (d <= kDistEps) return res;
    const double n[3] = {(s1.r[0] - s2.r[0]) / d, (s1.r[1] - s2.r[1]) / d, (s1.r[2] - s2.r[2]) / d};
    const double d0 = d * 1e-10;
    const double pref = kHbar * kMu0 / (4.0 * kPi) * (s1.iso.gamma * s2.iso.gamma) / (d0 * d0 * d0);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) res[3 * a + b] = pref * ((a == b ? 1.0 : 0.0) - 3.0 * n[a] * n[b]);
    return res;
}

// <state| S_a |state> for a spin-1 (or spin-1/2) source
std::array<double, 3> spin_vector(int mult, const cplx* st) {
    Mat sx, sy, sz;
    spin_ops(mult, sx, sy, sz);
    std::array<double, 3> out{};
    const Mat* ops[3] = {&sx, &sy, &sz};
    for (int a = 0; a < 3; ++a) {
        cplx acc = 0.0;
        for (int i = 0; i < mult; ++i)
            for (int j = 0; j < mult; ++j) acc += std::conj(st[i]) * (*ops[a])[i + mult * j] * st[j];
        out[a] = acc.real();
    }
    return out;
}

// Hermitian eigen-decomposition by cyclic complex Jacobi rotations;
// eigenvalues ascending, eigenvectors in the columns of V.
void eig_hermitian(int n, Mat A, std::vector<double>& w, Mat& V) {
    V.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 0; q < n; ++q)
            for (int p = 0; p < q; ++p) off += std::norm(A[p + n * q]);
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cplx apq = A[p + n * q];
                const double mag = std::abs(apq);
                if (mag == 0.0) continue;
                const double app = A[p + n * p].real(), aqq = A[q + n * q].real();
                const cplx ph = apq / mag;  // A_pq = mag * e^{i phi}
                const double theta = 0.5 * std::atan2(2.0 * mag, aqq - app);
                const double c = std::cos(theta), s = std::sin(theta);
                // rotation G acting on columns p, q: [c, s*ph; -s*conj(ph), c]
                for (int k = 0; k < n; ++k) {
                    const cplx akp = A[k + n * p], akq = A[k + n * q];
                    A[k + n * p] = c * akp - s * std::conj(ph) * akq;
                    A[k + n * q] = s * ph * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const cplx apk = A[p + n * k], aqk = A[q + n * k];
                    A[p + n * k] = c * apk - s * ph * aqk;
                    A[q + n * k] = s * std::conj(ph) * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const cplx vkp = V[k + n * p], vkq = V[k + n * q];
                    V[k + n * p] = c * vkp - s * std::conj(ph) * vkq;
                    V[k + n * q] = s * ph * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a + n * a].real() < A[b + n * b].real(); });
    w.resize(n);
    Mat Vs(n * n);
    for (int k = 0; k < n; ++k) {
        w[k] = A[idx[k] + n * idx[k]].real();
        for (int i = 0; i < n; ++i) Vs[i + n * k] = V[i + n * idx[k]];
    }
    V = Vs;
}

// ---------------------------------------------------------------------------
// Term lists
struct Term {
    double c;
    std::vector<int> idx;
    std::vector<Mat> mats;
};

struct Owned {
    int nspin = 0;
    std::vector<int64_t> spin_dim;
    std::vector<double> coeff;
    std::vector<int64_t> nbody, pos_off, pos, dims, mat_off;
    std::vector<double> mats;
    oops_term_table view() const {
        oops_term_table t;
        t.nspin = nspin;
        t.nterm = (int32_t)coeff.size();
        t.coeff = coeff.data();
        t.nbody = nbody.data();
        t.pos_offset = pos_off.data();
        t.pos = pos.data();
        t.dims = dims.data();
        t.mat_offset = mat_off.data();
        t.mats = mats.data();
        t.spin_dim = spin_dim.data();
        return t;
    }
};

Owned flatten(const std::vector<Term>& terms, const std::vector<int64_t>& spin_dim) {
    Owned o;
    o.nspin = (int)spin_dim.size();
    o.spin_dim = spin_dim;
    o.pos_off.push_back(0);
    o.mat_off.push_back(0);
    for (const Term& t : terms) {
        o.coeff.push_back(t.c);
        o.nbody.push_back((int64_t)t.idx.size());
        for (size_t b = 0; b < t.idx.size(); ++b) {
            o.pos.push_back(t.idx[b]);
            const int64_t d = (int64_t)std::llround(std::sqrt((double)t.mats[b].size()));
            o.dims.push_back(d);
            o.mat_off.push_back(o.mat_off.back() + d * d);
            for (const cplx& z : t.mats[b]) { o.mats.push_back(z.real()); o.mats.push_back(z.imag()); }
        }
        o.pos_off.push_back((int64_t)o.pos.size());
    }
    if (o.pos.empty()) { o.pos.push_back(0); o.dims.push_back(1); }
    if (o.mats.empty()) { o.mats.push_back(0.0); o.mats.push_back(0.0); }
    return o;
}

void add_filtered(std::vector<Term>& out, const std::vector<int>& idx, const double* coeffs,
                  const std::vector<std::vector<Mat>>& forms) {
    for (size_t k = 0; k < forms.size(); ++k)
        if (coeffs[k] != 0.0) out.push_back(Term{coeffs[k], idx, forms[k]});
}

std::vector<std::vector<Mat>> single_forms(int mult) {
    Mat sx, sy, sz;
    spin_ops(mult, sx, sy, sz);
    return {{sx}, {sy}, {sz}, {matmul(sx, sx, mult)}, {matmul(sy, sy, mult)}, {matmul(sz, sz, mult)}};
}

// create_spin_hamiltonian: dipolar + Zeeman + field of the centre spin, and
// for the single-sample variant the bath's mean field on each cluster spin
// (bath_field: 3 values per spin, single_sample_cce.cpp:75-93)
std::vector<Term> hamiltonian(const std::vector<SpinH>& spins, const double* B, const SpinH& center,
                              const cplx* cstate, const double* bath_field = nullptr) {
    std::vector<Term> out;
    const int k = (int)spins.size();
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j) {
            Mat a[3], b[3];
            spin_ops(spins[i].iso.mult, a[0], a[1], a[2]);
            spin_ops(spins[j].iso.mult, b[0], b[1], b[2]);
            std::vector<std::vector<Mat>> forms;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) forms.push_back({a[p], b[q]});
            const auto c = dipole(spins[i], spins[j]);
            add_filtered(out, {i, j}, c.data(), forms);
        }
    for (int i = 0; i < k; ++i) {
        const Isotope& I = spins[i].iso;
        const double c[6] = {-I.gamma * B[0], -I.gamma * B[1], -I.gamma * B[2], I.eta / 3.0, -I.eta / 3.0, I.omegaQ};
        add_filtered(out, {i}, c, single_forms(I.mult));
    }
    const auto sv = spin_vector(center.iso.mult, cstate);
    for (int i = 0; i < k; ++i) {
        const auto dp = dipole(spins[i], center);
        double f[6] = {0, 0, 0, 0, 0, 0};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) f[r] += dp[r + 3 * c] * sv[c];  // reshape(dip, 3, 3) * s
        add_filtered(out, {i}, f, single_forms(spins[i].iso.mult));
    }
    if (bath_field)
        for (int i = 0; i < k; ++i) {
            const double f[6] = {bath_field[3 * i], bath_field[3 * i + 1], bath_field[3 * i + 2], 0.0, 0.0, 0.0};
            add_filtered(out, {i}, f, single_forms(spins[i].iso.mult));
        }
    return out;
}

// glibc's rand() after srand(seed) (the additive feedback generator of
// random(3), TYPE_3: r_i = r_{i-3} + r_{i-31} mod 2^32 from a Park-Miller
// seeding of 31 words, 310 values discarded, outputs r_i >> 1), restated so
// that the bath states do not depend on or disturb the process's rand() state.
class GlibcRand {
public:
    explicit GlibcRand(uint32_t seed) {
        int32_t w = (int32_t)seed;
        if (w == 0) w = 1;
        r_[0] = w;
        for (int i = 1; i < 31; ++i) {
            const int64_t hi = r_[i - 1] / 127773, lo = r_[i - 1] % 127773;
            int32_t word = (int32_t)(16807 * lo - 2836 * hi);
            if (word < 0) word += 2147483647;
            r_[i] = word;
        }
        for (int i = 31; i < 34; ++i) r_[i] = r_[i - 31];
        n_ = 34;
        for (int i = 34; i < 344; ++i) step();
    }
    int32_t next() { return (int32_t)(step() >> 1); }

private:
    uint32_t step() {
        const uint32_t v = (uint32_t)r_[(n_ - 31) % 34] + (uint32_t)r_[(n_ - 3) % 34];
        r_[n_ % 34] = (int32_t)v;
        ++n_;
        return v;
    }
    int32_t r_[34];
    int64_t n_ = 0;
};

// Mean field of the bath on cluster spin `s`: sum over every bath spin j
// outside the cluster of dipole(s, j) <S_j>, with <S_j> = (0, 0, +1/2) for
// state 0 and (0, 0, -1/2) for state 1, in bath order
// (SpinInteraction.cpp:159-173, SpinInteractionComponent.cpp:433-451).
void bath_mean_field(const SpinH& s, const double* coords, int nspin, const int32_t* states,
                     const std::vector<int32_t>& cluster, double* f) {
    f[0] = f[1] = f[2] = 0.0;
    SpinH src;
    isotope_of("13C", src.iso);
    size_t q = 0;   // cluster is sorted ascending
    for (int j = 0; j < nspin; ++j) {
        while (q < cluster.size() && cluster[q] < j) ++q;
        if (q < cluster.size() && cluster[q] == j) continue;
        for (int c = 0; c < 3; ++c) src.r[c] = coords[3 * j + c];
        const auto dp = dipole(s, src);
        const double sz = states[j] == 0 ? 0.5 : -0.5;
        for (int r = 0; r < 3; ++r) f[r] += dp[r + 3 * 2] * sz;
    }
}

Mat kron(const Mat& A, int da, const Mat& B, int db) {
    const int d = da * db;
    Mat C(d * d);
    for (int ca = 0; ca < da; ++ca)
        for (int cb = 0; cb < db; ++cb)
            for (int ra = 0; ra < da; ++ra)
                for (int rb = 0; rb < db; ++rb)
                    C[(ra * db + rb) + d * (ca * db + cb)] = A[ra + da * ca] * B[rb + db * cb];
    return C;
}

Mat eye(int d) {
    Mat I(d * d, 0.0);
    for (int i = 0; i < d; ++i) I[i + d * i] = 1.0;
    return I;
}

Mat transpose(const Mat& A, int d) {
    Mat T(d * d);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) T[j + d * i] = A[i + d * j];
    return T;
}

Mat sharp(const Mat& m) { const int d = (int)std::llround(std::sqrt((double)m.size())); return kron(transpose(m, d), d, eye(d), d); }
Mat flat(const Mat& m) { const int d = (int)std::llround(std::sqrt((double)m.size())); return kron(eye(d), d, m, d); }

std::vector<Term> expand(const std::vector<Term>& h, bool use_sharp, double sign) {
    std::vector<Term> out;
    for (const Term& t : h) {
        Term e{t.c * sign, t.idx, {}};
        for (const Mat& m : t.mats) e.mats.push_back(use_sharp ? sharp(m) : flat(m));
        out.push_back(e);
    }
    return out;
}

// i (sigma^T (x) sigma - I) of the dephasing axis (SpinInteractionComponent.cpp:285-304)
Mat dephasing_matrix(const double* axis) {
    double nrm = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (nrm == 0.0) nrm = 1.0;
    const double nx = axis[0] / nrm, ny = axis[1] / nrm, nz = axis[2] / nrm;
    const Mat sig = {nz, cplx(nx, ny), cplx(nx, -ny), -nz};  // [[nz, nx-i ny],[nx+i ny, -nz]] column-major
    Mat dm = kron(transpose(sig, 2), 2, sig, 2);
    for (int i = 0; i < 4; ++i) dm[i + 4 * i] -= 1.0;
    for (auto& z : dm) z *= cplx(0.0, 1.0);
    return dm;
}

std::vector<Term> dephasing(int k, double rate_hz, const double* axis) {
    std::vector<Term> out;
    const double rate = 2.0 * kPi * rate_hz;
    const Mat dm = dephasing_matrix(axis);
    for (int i = 0; i < k; ++i) {
        const double c[1] = {rate};
        add_filtered(out, {i}, c, {{dm}});
    }
    return out;
}

std::vector<Term> negated_adjoint(const std::vector<Term>& terms) {
    std::vector<Term> out;
    for (const Term& t : terms) {
        Term e{-t.c, t.idx, {}};
        for (const Mat& m : t.mats) {
            const int d = (int)std::llround(std::sqrt((double)m.size()));
            Mat a(d * d);
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j) a[j + d * i] = std::conj(m[i + d * j]);
            e.mats.push_back(a);
        }
        out.push_back(e);
    }
    return out;
}

// vec(rho0) = vecterize(sum pol.S_i + I/D) in the per-spin block order
std::vector<cplx> density_vector(const std::vector<SpinH>& spins, const double* pol, double& D) {
    const int k = (int)spins.size();
    std::vector<Term> terms;
    for (int i = 0; i < k; ++i) {
        const double c[6] = {pol[0], pol[1], pol[2], 0.0, 0.0, 0.0};
        add_filtered(terms, {i}, c, single_forms(spins[i].iso.mult));
    }
    D = 1.0;
    for (const auto& s : spins) D *= s.iso.mult;
    terms.push_back(Term{1.0 / D, {}, {}});
    int64_t n = 1;
    for (const auto& s : spins) n *= (int64_t)s.iso.mult * s.iso.mult;
    std::vector<cplx> v(n, 0.0);
    for (const Term& t : terms) {
        std::vector<cplx> acc(1, 1.0);
        for (int s = 0; s < k; ++s) {
            const int d = spins[s].iso.mult;
            Mat m = eye(d);
            for (size_t b = 0; b < t.idx.size(); ++b)
                if (t.idx[b] == s) m = t.mats[b];
            std::vector<cplx> nxt(acc.size() * d * d);
            for (size_t a = 0; a < acc.size(); ++a)
                for (int e = 0; e < d * d; ++e) nxt[a * d * d + e] = acc[a] * m[e];
            acc.swap(nxt);
        }
        for (int64_t i = 0; i < n; ++i) v[i] += t.c * acc[i];
    }
    return v;
}

// ---------------------------------------------------------------------------
struct NV {
    SpinH espin;
    std::vector<double> evals;
    Mat evecs;
};

NV make_nv(const double* pos, const double* B) {
    NV nv;
    const double d4 = kNVLattice / 4.0;
    nv.espin.r[0] = pos[0] + d4;
    nv.espin.r[1] = pos[1] + d4;
    nv.espin.r[2] = pos[2] + d4;
    isotope_of("NVe", nv.espin.iso);
    Mat sx, sy, sz;
    spin_ops(3, sx, sy, sz);
    const double omegaQ = nv.espin.iso.omegaQ * 2.0 * kPi * 1e6;
    const double g = nv.espin.iso.gamma;
    Mat s111(9);
    for (int i = 0; i < 9; ++i) s111[i] = (sx[i] + sy[i] + sz[i]) / std::sqrt(3.0);
    Mat s2 = matmul(s111, s111, 3);
    Mat H(9);
    for (int i = 0; i < 9; ++i) H[i] = omegaQ * s2[i] + g * (B[0] * sx[i] + B[1] * sy[i] + B[2] * sz[i]);
    eig_hermitian(3, H, nv.evals, nv.evecs);
    return nv;
}

struct ClusterProblem {
    Owned A, Bd;          // lvA and -lvB^dagger (Hahn-echo sweeps)
    Owned lvB, Ad;        // lvB and -lvA^dagger (CPMG-N segments)
    Owned H0, H1;         // for inspection
    std::vector<cplx> rho;
    double D = 1.0;
};

enum : int { kNeedEcho = 1, kNeedCpmg = 2, kNeedHam = 4, kNeedAll = 7 };

ClusterProblem build_problem(const std::vector<SpinH>& spins, const oops_cce_params& P, const NV& nv,
                             int need = kNeedAll) {
    ClusterProblem cp;
    const cplx* st0 = &nv.evecs[3 * P.state0];
    const cplx* st1 = &nv.evecs[3 * P.state1];
    auto h0 = hamiltonian(spins, P.B, nv.espin, st0);
    auto h1 = hamiltonian(spins, P.B, nv.espin, st1);
    auto deph = dephasing((int)spins.size(), P.dephasing_rate, P.dephasing_axis);
    auto lvA = expand(h0, true, 1.0);
    auto f1 = expand(h1, false, -1.0);
    lvA.insert(lvA.end(), f1.begin(), f1.end());
    lvA.insert(lvA.end(), deph.begin(), deph.end());
    auto lvB = expand(h1, true, 1.0);
    auto f0 = expand(h0, false, -1.0);
    lvB.insert(lvB.end(), f0.begin(), f0.end());
    lvB.insert(lvB.end(), deph.begin(), deph.end());
    std::vector<int64_t> hd, ldim;
    for (const auto& s : spins) { hd.push_back(s.iso.mult); ldim.push_back((int64_t)s.iso.mult * s.iso.mult); }
    if (need & kNeedHam) {
        cp.H0 = flatten(h0, hd);
        cp.H1 = flatten(h1, hd);
    }
    cp.A = flatten(lvA, ldim);
    cp.Bd = flatten(negated_adjoint(lvB), ldim);
    if (need & (kNeedCpmg | kNeedHam)) {
        cp.lvB = flatten(lvB, ldim);
        cp.Ad = flatten(negated_adjoint(lvA), ldim);
    }
    cp.rho = density_vector(spins, P.pol, cp.D);
    return cp;
}

// SingleSampleCCE::cluster_evolution (single_sample_cce.cpp:48-73, 75-105):
// the Hilbert-space conditional Hamiltonians with the bath mean field and the
// product state of the cluster's spins.
struct SingleSampleProblem {
    Owned H0, H1;
    std::vector<cplx> psi;
};

SingleSampleProblem build_single_sample(const double* coords, int nspin, const std::vector<int32_t>& cluster,
                                        const int32_t* states, const oops_cce_params& P, const NV& nv) {
    const int k = (int)cluster.size();
    std::vector<SpinH> spins(k);
    std::vector<double> bf(3 * (size_t)k);
    for (int i = 0; i < k; ++i) {
        for (int c = 0; c < 3; ++c) spins[i].r[c] = coords[3 * cluster[i] + c];
        isotope_of("13C", spins[i].iso);
    }
    for (int i = 0; i < k; ++i) bath_mean_field(spins[i], coords, nspin, states, cluster, &bf[3 * i]);
    SingleSampleProblem sp;
    std::vector<int64_t> hd(k, 2);
    sp.H0 = flatten(hamiltonian(spins, P.B, nv.espin, &nv.evecs[3 * P.state0], bf.data()), hd);
    sp.H1 = flatten(hamiltonian(spins, P.B, nv.espin, &nv.evecs[3 * P.state1], bf.data()), hd);
    // kron of the spins' basis vectors, spin 0 the slowest factor
    int64_t idx = 0;
    for (int i = 0; i < k; ++i) idx = 2 * idx + (states[cluster[i]] ? 1 : 0);
    sp.psi.assign((size_t)1 << k, 0.0);
    sp.psi[idx] = 1.0;
    return sp;
}

// Constants of the device term builder: Zeeman coefficients, NV position and
// state spin vectors, dephasing superoperator, SHARP / FLAT of the spin-1/2
// operators (computed by the same host functions as the host path), and
// rho0 = vec(I / 2^k) for k = 1..4.
void builder_consts(const oops_cce_params& P, const NV& nv, oops::TermBuilderArgs& a,
                    std::vector<std::vector<double>>& rho) {
    std::memset(&a, 0, sizeof(a));
    Isotope c13;
    isotope_of("13C", c13);
    for (int c = 0; c < 3; ++c) {
        a.zee[c] = -c13.gamma * P.B[c];
        a.center[c] = nv.espin.r[c];
    }
    a.gamma_c = c13.gamma;
    a.gamma_e = nv.espin.iso.gamma;
    const auto s0 = spin_vector(nv.espin.iso.mult, &nv.evecs[3 * P.state0]);
    const auto s1 = spin_vector(nv.espin.iso.mult, &nv.evecs[3 * P.state1]);
    for (int c = 0; c < 3; ++c) {
        a.sv0[c] = s0[c];
        a.sv1[c] = s1[c];
    }
    a.deph_rate = 2.0 * kPi * P.dephasing_rate;
    const Mat dm = dephasing_matrix(P.dephasing_axis);
    for (int q = 0; q < 16; ++q) a.dm[q] = make_double2(dm[q].real(), dm[q].imag());
    Mat sx, sy, sz;
    spin_ops(2, sx, sy, sz);
    const Mat* ops[3] = {&sx, &sy, &sz};
    for (int c = 0; c < 3; ++c) {
        const Mat sh = sharp(*ops[c]), fl = flat(*ops[c]);
        for (int q = 0; q < 16; ++q) {
            a.sharp[16 * c + q] = make_double2(sh[q].real(), sh[q].imag());
            a.flat[16 * c + q] = make_double2(fl[q].real(), fl[q].imag());
        }
    }
    rho.assign(5, {});
    const double zero[3] = {0.0, 0.0, 0.0};
    for (int k = 1; k <= 4; ++k) {
        std::vector<SpinH> spins(k);
        for (auto& sp : spins) {
            isotope_of("13C", sp.iso);
            sp.r[0] = sp.r[1] = sp.r[2] = 0.0;
        }
        double D = 1.0;
        const auto v = density_vector(spins, zero, D);
        for (const cplx& z : v) { rho[k].push_back(z.real()); rho[k].push_back(z.imag()); }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
struct oops_terms {
    Owned o;
};

struct oops_clusters {
    int nspin = 0;
    int max_order = 0;
    std::vector<std::vector<std::vector<int32_t>>> idx;      // [order][cluster] -> sorted spins
    std::vector<std::vector<std::vector<int32_t>>> parents;  // [order][cluster] -> parent positions
};

// The device term builder takes the 13C clusters of at most four spins when
// the bath is unpolarised (rho0 = I / D for every cluster); OOPS_TERM_BUILDER
// = host keeps the host compiler.
static bool device_builder_applies(const oops_clusters* C, const oops_cce_params* P) {
    if (const char* e = std::getenv("OOPS_TERM_BUILDER"))
        if (std::string(e) == "host") return false;
    if (P->pol[0] != 0.0 || P->pol[1] != 0.0 || P->pol[2] != 0.0) return false;
    return C->max_order <= 4 && C->nspin > 0;
}


// Selected clusters grouped by size for the device term builder; items are
// numbered in selection order (global cluster order).
static std::vector<oops::BuiltGroup> builder_groups(const oops_clusters* C, int64_t nsel, const int64_t* sel, int64_t& nitems,
                                             bool& bad_selection) {
    std::vector<oops::BuiltGroup> groups(C->max_order);
    int64_t g = 0, next = 0;
    nitems = 0;
    for (int o = 0; o < C->max_order; ++o) {
        groups[o].k = o + 1;
        for (const auto& cl : C->idx[o]) {
            const int64_t gi = g++;
            if (sel) {
                if (next >= nsel || sel[next] != gi) continue;
                ++next;
            }
            groups[o].spins.insert(groups[o].spins.end(), cl.begin(), cl.end());
            groups[o].items.push_back(nitems++);
        }
    }
    bad_selection = sel && next != nsel;
    return groups;
}


extern "C" {

int oops_read_xyz(const char* path, int32_t max_n, double* coords, char* isotopes, int32_t* n_out) {
    if (!path || !n_out || max_n < 0 || (max_n > 0 && !coords)) return OOPS_ERR_ARG;
    std::ifstream f(path);
    if (!f) return OOPS_ERR_ARG;
    long long nb = -1;
    if (!(f >> nb) || nb < 0) return OOPS_ERR_ARG;
    *n_out = (int32_t)nb;
    if (nb > max_n) return OOPS_ERR_CAPACITY;
    for (long long i = 0; i < nb; ++i) {
        std::string iso;
        double x, y, z;
        if (!(f >> iso >> x >> y >> z)) return OOPS_ERR_ARG;
        coords[3 * i] = x;
        coords[3 * i + 1] = y;
        coords[3 * i + 2] = z;
        if (isotopes) {
            char* slot = isotopes + 16 * i;
            std::memset(slot, 0, 16);
            std::memcpy(slot, iso.data(), std::min<size_t>(iso.size(), 15));
        }
    }
    return OOPS_OK;
}

int oops_write_xyz(const char* path, int32_t n, const double* coords, const char* isotopes) {
    if (!path || n < 0 || (n > 0 && !coords)) return OOPS_ERR_ARG;
    std::ofstream f(path);
    if (!f) return OOPS_ERR_ARG;
    f << n << std::endl;
    for (int32_t i = 0; i < n; ++i) {
        f << (isotopes ? std::string(isotopes + 16 * i, strnlen(isotopes + 16 * i, 16)) : std::string("13C")) << "\t";
        f << coords[3 * i] << "\t" << coords[3 * i + 1] << "\t" << coords[3 * i + 2] << std::endl;
    }
    return f ? OOPS_OK : OOPS_ERR_ARG;
}

int oops_bath_diamond(double a, double radius, double p, uint64_t seed, int32_t shape, double* coords,
                      int32_t max_n, int32_t* n_out) {
    if (!(a > 0.0) || !(radius > 0.0) || p < 0.0 || p > 1.0 || !n_out) return OOPS_ERR_ARG;
    static const double basis[8][3] = {{0, 0, 0}, {0, .5, .5}, {.5, 0, .5}, {.5, .5, 0},
                                       {.25, .25, .25}, {.25, .75, .75}, {.75, .25, .75}, {.75, .75, .25}};
    const int N = (int)std::ceil(radius / a) + 1;
    uint64_t state = seed;
    auto next_u = [&]() {  // splitmix64 -> uniform [0,1)
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    int32_t cnt = 0;
    for (int i = -N; i <= N; ++i)
        for (int j = -N; j <= N; ++j)
            for (int k = -N; k <= N; ++k)
                for (int b = 0; b < 8; ++b) {
                    const double x = a * (i + basis[b][0]), y = a * (j + basis[b][1]), z = a * (k + basis[b][2]);
                    const bool nsite = (i == 0 && j == 0 && k == 0 && (b == 0 || b == 4));  // N and V of the NV
                    if (nsite) continue;
                    const bool inside = shape == 0 ? (x * x + y * y + z * z <= radius * radius)
                                                   : (std::fabs(x) <= radius && std::fabs(y) <= radius && std::fabs(z) <= radius);
                    if (!inside) continue;
                    if (next_u() < p) {
                        if (coords && cnt < max_n) {
                            coords[3 * cnt] = x;
                            coords[3 * cnt + 1] = y;
                            coords[3 * cnt + 2] = z;
                        }
                        ++cnt;
                    }
                }
    *n_out = cnt;
    if (coords && cnt > max_n) return OOPS_ERR_CAPACITY;
    return OOPS_OK;
}

int oops_nv_states(const double* pos, const double* B, double* evals, double* evecs) {
    if (!pos || !B) return OOPS_ERR_ARG;
    NV nv = make_nv(pos, B);
    if (evals) for (int i = 0; i < 3; ++i) evals[i] = nv.evals[i];
    if (evecs) for (int i = 0; i < 9; ++i) { evecs[2 * i] = nv.evecs[i].real(); evecs[2 * i + 1] = nv.evecs[i].imag(); }
    return OOPS_OK;
}

int oops_cce_cluster_terms(const double* coords, int32_t k, const oops_cce_params* P, int32_t which,
                           oops_terms** out, double* rho, double* D) {
    if (!coords || k <= 0 || !P || !out) return OOPS_ERR_ARG;
    *out = nullptr;
    std::vector<SpinH> spins(k);
    for (int i = 0; i < k; ++i) {
        for (int c = 0; c < 3; ++c) spins[i].r[c] = coords[3 * i + c];
        isotope_of("13C", spins[i].iso);
    }
    NV nv = make_nv(P->nv_pos, P->B);
    ClusterProblem cp = build_problem(spins, *P, nv);
    std::unique_ptr<oops_terms> t(new oops_terms());
    switch (which) {
        case 0: t->o = cp.A; break;
        case 1: t->o = cp.lvB; break;
        case 2: t->o = cp.Bd; break;
        case 3: t->o = cp.H0; break;
        case 4: t->o = cp.H1; break;
        case 5: t->o = cp.Ad; break;
        default: return OOPS_ERR_ARG;
    }
    if (rho) for (size_t i = 0; i < cp.rho.size(); ++i) { rho[2 * i] = cp.rho[i].real(); rho[2 * i + 1] = cp.rho[i].imag(); }
    if (D) *D = cp.D;
    *out = t.release();
    return OOPS_OK;
}

int oops_terms_view(const oops_terms* t, oops_term_table* view) {
    if (!t || !view) return OOPS_ERR_ARG;
    *view = t->o.view();
    return OOPS_OK;
}

int oops_terms_sizes(const oops_terms* t, int64_t* nterm, int64_t* npos, int64_t* nmat) {
    if (!t) return OOPS_ERR_ARG;
    if (nterm) *nterm = (int64_t)t->o.coeff.size();
    if (npos) *npos = t->o.pos_off.back();
    if (nmat) *nmat = t->o.mat_off.back();
    return OOPS_OK;
}

int oops_terms_destroy(oops_terms* t) {
    delete t;
    return OOPS_OK;
}

}  // extern "C"

// Connected vertex sets of 1..max_order vertices of the graph with the sorted
// neighbour lists nbr, in the reference's set order with parent lists
// (cDepthFirstPathTracing::generate, SpinClusterAlgorithm.cpp:98-148).
static void enumerate_graph(const std::vector<std::vector<int32_t>>& nbr, int nspin, int max_order,
                            oops_clusters* C) {
    C->nspin = nspin;
    C->max_order = max_order;
    C->idx.resize(max_order);
    C->parents.resize(max_order);
    for (int i = 0; i < nspin; ++i) {
        C->idx[0].push_back({i});
        C->parents[0].push_back({});
    }
    std::vector<char> inpar(nspin, 0), cand(nspin, 0);
    for (int order = 1; order < max_order; ++order) {
        std::map<std::vector<int32_t>, std::vector<int32_t>> gen;  // lexicographic == reference set order
        const auto& prev = C->idx[order - 1];
        std::vector<int32_t> cl;
        for (int pi = 0; pi < (int)prev.size(); ++pi) {
            const auto& par = prev[pi];
            for (int v : par) inpar[v] = 1;
            cl.clear();
            for (int v : par)
                for (int u : nbr[v])
                    if (!inpar[u] && !cand[u]) { cand[u] = 1; cl.push_back(u); }
            std::sort(cl.begin(), cl.end());
            for (int u : cl) {
                std::vector<int32_t> child(par);
                child.insert(std::upper_bound(child.begin(), child.end(), u), u);
                gen[child].push_back(pi);
                cand[u] = 0;
            }
            for (int v : par) inpar[v] = 0;
        }
        for (auto& kv : gen) {
            C->idx[order].push_back(kv.first);
            C->parents[order].push_back(kv.second);
        }
    }
}

extern "C" {

int oops_clusters_enumerate(const double* coords, int32_t nspin, double cutoff, int32_t max_order,
                            oops_clusters** out) {
    if (!out || nspin < 0 || max_order < 1 || max_order > 10 || (nspin > 0 && !coords)) return OOPS_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<oops_clusters> C(new oops_clusters());
    // connection lists: d(i,j) computed for j >= i and mirrored (SpinCollection.cpp:38-47)
    std::vector<std::vector<int32_t>> nbr(nspin);
    for (int i = 0; i < nspin; ++i)
        for (int j = i + 1; j < nspin; ++j)
            if (dist3(coords + 3 * i, coords + 3 * j) <= cutoff) {
                nbr[i].push_back(j);
                nbr[j].push_back(i);
            }
    for (auto& v : nbr) std::sort(v.begin(), v.end());
    enumerate_graph(nbr, nspin, max_order, C.get());
    *out = C.release();
    return OOPS_OK;
}

int oops_clusters_enumerate_graph(int32_t nspin, const int64_t* offsets, const int32_t* neighbors,
                                  int32_t max_order, oops_clusters** out) {
    if (!out || nspin < 0 || max_order < 1 || max_order > 10 || (nspin > 0 && !offsets)) return OOPS_ERR_ARG;
    *out = nullptr;
    std::vector<std::vector<int32_t>> nbr(nspin);
    for (int i = 0; i < nspin; ++i) {
        if (offsets[i + 1] < offsets[i] || (offsets[i + 1] > offsets[i] && !neighbors)) return OOPS_ERR_ARG;
        for (int64_t q = offsets[i]; q < offsets[i + 1]; ++q) {
            const int32_t j = neighbors[q];
            if (j < 0 || j >= nspin) return OOPS_ERR_ARG;
            if (j != i) nbr[i].push_back(j);   // the connection matrix's diagonal is zeroed
        }
    }
    // symmetric: an edge listed on either side connects both (getConnectionMatrix mirrors j >= i)
    for (int i = 0; i < nspin; ++i)
        for (int32_t j : std::vector<int32_t>(nbr[i])) nbr[j].push_back(i);
    for (auto& v : nbr) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::unique_ptr<oops_clusters> C(new oops_clusters());
    enumerate_graph(nbr, nspin, max_order, C.get());
    *out = C.release();
    return OOPS_OK;
}

int oops_clusters_count(const oops_clusters* C, int32_t order, int64_t* count) {
    if (!C || !count || order < 0 || order >= C->max_order) return OOPS_ERR_ARG;
    *count = (int64_t)C->idx[order].size();
    return OOPS_OK;
}

int oops_clusters_get(const oops_clusters* C, int32_t order, int32_t* idx) {
    if (!C || !idx || order < 0 || order >= C->max_order) return OOPS_ERR_ARG;
    size_t q = 0;
    for (const auto& c : C->idx[order])
        for (int32_t v : c) idx[q++] = v;
    return OOPS_OK;
}

int oops_clusters_parents(const oops_clusters* C, int32_t order, int64_t* offsets, int32_t* list) {
    if (!C || order < 0 || order >= C->max_order) return OOPS_ERR_ARG;
    int64_t q = 0;
    const auto& P = C->parents[order];
    for (size_t i = 0; i < P.size(); ++i) {
        if (offsets) offsets[i] = q;
        for (int32_t p : P[i]) {
            if (list) list[q] = p;
            ++q;
        }
    }
    if (offsets) offsets[P.size()] = q;
    return OOPS_OK;
}

int oops_clusters_destroy(oops_clusters* C) {
    delete C;
    return OOPS_OK;
}

// L holds every cluster's L_c(t) row by row, orders concatenated in order.
int oops_cce_reduce(const oops_clusters* C, int32_t nt, const double* L, double* tilde, double* final_result,
                    double* each_order, double* log_abs_final, int32_t* sign_final) {
    if (!C || nt <= 0 || !L) return OOPS_ERR_ARG;
    const int K = C->max_order;
    std::vector<int64_t> base(K + 1, 0);
    for (int o = 0; o < K; ++o) base[o + 1] = base[o] + (int64_t)C->idx[o].size();
    std::vector<double> til((size_t)base[K] * nt);
    std::vector<double> ltil((size_t)base[K] * nt);
    std::vector<int8_t> stil((size_t)base[K] * nt);
    // memoised sub-cluster sets (getSubClusters) as flat global indices, sorted by (order, position)
    std::vector<std::vector<int64_t>> subs(base[K]);
    for (int o = 0; o < K; ++o)
        for (int64_t j = 0; j < (int64_t)C->idx[o].size(); ++j) {
            std::set<int64_t> s;
            if (o > 0)
                for (int32_t p : C->parents[o][j]) {
                    const int64_t g = base[o - 1] + p;
                    s.insert(g);
                    for (int64_t q : subs[g]) s.insert(q);
                }
            subs[base[o] + j].assign(s.begin(), s.end());
        }
    for (int o = 0; o < K; ++o)
        for (int64_t j = 0; j < (int64_t)C->idx[o].size(); ++j) {
            const int64_t g = base[o] + j;
            for (int t = 0; t < nt; ++t) {
                double r = L[g * nt + t];
                double lr = std::log(std::fabs(r));
                int sg = r < 0 ? -1 : 1;
                for (int64_t q : subs[g]) {
                    r = r / til[q * nt + t];
                    lr -= ltil[q * nt + t];
                    sg *= stil[q * nt + t];
                }
                til[g * nt + t] = r;
                ltil[g * nt + t] = lr;
                stil[g * nt + t] = (int8_t)sg;
            }
        }
    if (tilde) std::memcpy(tilde, til.data(), sizeof(double) * til.size());
    std::vector<double> acc(nt, 1.0), lacc(nt, 0.0);
    std::vector<int> sacc(nt, 1);
    for (int o = 0; o < K; ++o) {
        std::vector<double> r(nt, 1.0), lr(nt, 0.0);
        std::vector<int> sr(nt, 1);
        for (int64_t j = 0; j < (int64_t)C->idx[o].size(); ++j) {
            const int64_t g = base[o] + j;
            for (int t = 0; t < nt; ++t) {
                r[t] = r[t] * til[g * nt + t];
                lr[t] += ltil[g * nt + t];
                sr[t] *= stil[g * nt + t];
            }
        }
        for (int t = 0; t < nt; ++t) {
            acc[t] = acc[t] * r[t];
            lacc[t] += lr[t];
            sacc[t] *= sr[t];
            if (each_order) each_order[(size_t)t * K + o] = r[t];
            if (final_result) final_result[(size_t)t * K + o] = acc[t];
            if (log_abs_final) log_abs_final[(size_t)t * K + o] = lacc[t];
            if (sign_final) sign_final[(size_t)t * K + o] = sacc[t];
        }
    }
    return OOPS_OK;
}

// Build every cluster's Hahn-echo problem (orders in sequence, clusters in
// set order) and hand the batch to the device runtime.
int oops_cce_hahn_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C, const oops_cce_params* P,
                               int64_t nsel, const int64_t* sel, int32_t m, int32_t klim, double tol,
                               int32_t precision, oops_hahn_batch** out) {
    if (!ctx || !coords || !C || !P || !out || nsel < 0 || (nsel > 0 && !sel)) return OOPS_ERR_ARG;
    NV nv = make_nv(P->nv_pos, P->B);
    if (device_builder_applies(C, P)) {
        int64_t nitems = 0;
        bool bad = false;
        const auto groups = builder_groups(C, nsel, sel, nitems, bad);
        if (bad) return oops::set_err(ctx, OOPS_ERR_ARG, "selection must be ascending cluster indices");
        oops::TermBuilderArgs consts;
        std::vector<std::vector<double>> rho;
        builder_consts(*P, nv, consts, rho);
        return oops::batch_create_built(ctx, coords, C->nspin, groups, nitems, consts, rho, 1, false, 0, nullptr, m,
                                        klim, tol, precision, out);
    }
    std::vector<ClusterProblem> probs;
    int64_t g = 0, next = 0;  // global cluster index, next selection
    for (int o = 0; o < C->max_order; ++o)
        for (const auto& cl : C->idx[o]) {
            const int64_t gi = g++;
            if (sel) {
                if (next >= nsel || sel[next] != gi) continue;
                ++next;
            }
            std::vector<SpinH> spins(cl.size());
            for (size_t i = 0; i < cl.size(); ++i) {
                for (int c = 0; c < 3; ++c) spins[i].r[c] = coords[3 * cl[i] + c];
                isotope_of("13C", spins[i].iso);
            }
            probs.push_back(build_problem(spins, *P, nv, kNeedEcho));
        }
    if (sel && next != nsel) return oops::set_err(ctx, OOPS_ERR_ARG, "selection must be ascending cluster indices");
    std::vector<oops_term_table> views(2 * probs.size());
    std::vector<oops_hahn_item> items(probs.size());
    for (size_t i = 0; i < probs.size(); ++i) {
        views[2 * i] = probs[i].A.view();
        views[2 * i + 1] = probs[i].Bd.view();
        items[i].opA = &views[2 * i];
        items[i].opB = &views[2 * i + 1];
        items[i].u0 = reinterpret_cast<const double*>(probs[i].rho.data());
        items[i].y0 = reinterpret_cast<const double*>(probs[i].rho.data());
        items[i].scale = probs[i].D;
    }
    return oops_hahn_batch_create(ctx, (int32_t)items.size(), items.data(), m, klim, tol, precision, out);
}

int oops_cce_echo_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C, const oops_cce_params* P,
                               int32_t npulse, int32_t nt, const double* t, int32_t m, int32_t klim, double tol,
                               int32_t precision, oops_hahn_batch** out) {
    if (!ctx || !coords || !C || !P || !out || nt < 1 || !t || npulse < 1) return OOPS_ERR_ARG;
    if (device_builder_applies(C, P)) {
        int64_t nitems = 0;
        bool bad = false;
        const auto groups = builder_groups(C, 0, nullptr, nitems, bad);
        NV nv = make_nv(P->nv_pos, P->B);
        oops::TermBuilderArgs consts;
        std::vector<std::vector<double>> rho;
        builder_consts(*P, nv, consts, rho);
        return oops::batch_create_built(ctx, coords, C->nspin, groups, nitems, consts, rho, npulse, true, nt, t, m,
                                        klim, tol, precision, out);
    }
    oops::CceItems it;
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    int rc = oops::cce_build_items(coords, C, P, npulse, nt, t, threads, it);
    if (rc) return oops::set_err(ctx, rc, "building cluster operators failed");
    return oops::batch_create(ctx, (int32_t)it.items.size(), it.items.data(), true, m, klim, tol, precision, out);
}

int oops_cce_op_tables(oops_ctx* ctx, const double* coords, int32_t k, const oops_cce_params* P, int32_t which,
                       int32_t device, int32_t max_loc, int32_t max_ent, int32_t* counts, int32_t* locs, double* val,
                       int32_t* off, uint8_t* eslot, int32_t* masks) {
    if (!ctx || !coords || k < 1 || !P || which < 0 || which > 3 || !counts) return OOPS_ERR_ARG;
    std::vector<oops::DevOpHeader> hdr(1);
    std::vector<oops::DevOpLocal> lv;
    std::vector<double> vv;
    std::vector<int32_t> ov, mv(oops::kSlotMaskStride, 0), sm(oops::kTBSummary, 0);
    std::vector<uint8_t> ev;
    NV nv = make_nv(P->nv_pos, P->B);
    if (!device) {
        std::vector<SpinH> spins(k);
        for (int i = 0; i < k; ++i) {
            for (int c = 0; c < 3; ++c) spins[i].r[c] = coords[3 * i + c];
            isotope_of("13C", spins[i].iso);
        }
        ClusterProblem cp = build_problem(spins, *P, nv);
        const Owned& o = which == 0 ? cp.A : (which == 1 ? cp.Bd : (which == 2 ? cp.lvB : cp.Ad));
        const oops_term_table T = o.view();
        oops::CompiledOp C;
        std::string err;
        int rc = oops::compile_operator(&T, C, err, 1 << 20);
        if (rc) return oops::set_err(ctx, rc, err);
        lv = C.locs;
        for (const auto& z : C.val) { vv.push_back(z.real()); vv.push_back(z.imag()); }
        ov = C.off;
        ev = C.eslot;
        for (int q = 0; q < oops::kSlotMaskStride; ++q) {
            if (q < oops::kSlot64 && q < (int)C.mask64.size()) mv[q] = C.mask64[q];
            if (q >= oops::kSlot64 && q - oops::kSlot64 < oops::kSlot32 && q - oops::kSlot64 < (int)C.mask32.size())
                mv[q] = C.mask32[q - oops::kSlot64];
        }
        sm = {(int32_t)C.locs.size(), (int32_t)C.val.size(), C.max_row, C.nslot64, C.nslot32, C.hermitian ? 1 : 0, 0,
              0};
    } else {
        if (k > 4) return oops::set_err(ctx, OOPS_ERR_UNSUPPORTED, "the device term builder takes 1-4 spins");
        if (hipSetDevice(ctx->device) != hipSuccess) return oops::set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
        oops::TermBuilderArgs a;
        std::vector<std::vector<double>> rho;
        builder_consts(*P, nv, a, rho);
        const int P2 = k * (k - 1) / 2, ls = k + P2, es = 16 * k + 128 * P2, nops = 4;
        std::vector<int32_t> cl(k);
        for (int i = 0; i < k; ++i) cl[i] = i;
        oops::DevBuf dc, dcl, dh, dl, dv, doff, de, dsm, dmm, dsum;
        auto al = [&](oops::DevBuf& b, size_t bytes) {
            const int rc = oops::hip_check(ctx, hipMalloc(&b.p, bytes), "hipMalloc");
            if (!rc) b.bytes = bytes;
            return rc;
        };
        int rc = al(dc, sizeof(double) * 3 * k);
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(dc.p, coords, sizeof(double) * 3 * k, hipMemcpyHostToDevice), "H2D");
        rc = rc ? rc : al(dcl, sizeof(int32_t) * k);
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(dcl.p, cl.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice), "H2D");
        rc = rc ? rc : al(dh, sizeof(oops::DevOpHeader) * nops);
        rc = rc ? rc : al(dl, sizeof(oops::DevOpLocal) * nops * ls);
        rc = rc ? rc : al(dv, sizeof(double2) * nops * es);
        rc = rc ? rc : al(doff, sizeof(int32_t) * nops * es);
        rc = rc ? rc : al(de, (size_t)nops * es);
        rc = rc ? rc : al(dsm, sizeof(int32_t) * nops * oops::kSlotMaskStride);
        rc = rc ? rc : al(dmm, sizeof(int32_t) * nops * oops::kMedMaskStride);
        rc = rc ? rc : al(dsum, sizeof(int32_t) * nops * oops::kTBSummary);
        if (rc) return rc;
        a.coords = (const double*)dc.p;
        a.clusters = (const int32_t*)dcl.p;
        a.nitems = 1;
        a.k = k;
        a.nops = nops;
        a.hdr = (oops::DevOpHeader*)dh.p;
        a.loc = (oops::DevOpLocal*)dl.p;
        a.val = (double2*)dv.p;
        a.off = (int32_t*)doff.p;
        a.eslot = (uint8_t*)de.p;
        a.smask = (int32_t*)dsm.p;
        a.mmask = (int32_t*)dmm.p;
        a.summary = (int32_t*)dsum.p;
        a.loc_stride = ls;
        a.ent_stride = es;
        if (oops::launch_term_builder(a, nops, ctx->stream) != 0)
            return oops::set_err(ctx, OOPS_ERR_HIP, "term builder launch failed");
        rc = oops::hip_check(ctx, hipStreamSynchronize(ctx->stream), "term builder");
        if (rc) return rc;
        rc = oops::hip_check(ctx, hipMemcpy(sm.data(), (int32_t*)dsum.p + which * oops::kTBSummary,
                                            sizeof(int32_t) * oops::kTBSummary, hipMemcpyDeviceToHost), "D2H");
        if (rc) return rc;
        lv.resize(sm[0]);
        vv.resize(2 * (size_t)sm[1]);
        ov.resize(sm[1]);
        ev.resize(sm[1]);
        rc = oops::hip_check(ctx, hipMemcpy(lv.data(), (oops::DevOpLocal*)dl.p + (size_t)which * ls, sizeof(oops::DevOpLocal) * lv.size(),
                                            hipMemcpyDeviceToHost), "D2H");
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(vv.data(), (double2*)dv.p + (size_t)which * es,
                                                      sizeof(double) * vv.size(), hipMemcpyDeviceToHost), "D2H");
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(ov.data(), (int32_t*)doff.p + (size_t)which * es,
                                                      sizeof(int32_t) * ov.size(), hipMemcpyDeviceToHost), "D2H");
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(ev.data(), (uint8_t*)de.p + (size_t)which * es, ev.size(),
                                                      hipMemcpyDeviceToHost), "D2H");
        rc = rc ? rc : oops::hip_check(ctx, hipMemcpy(mv.data(), (int32_t*)dsm.p + which * oops::kSlotMaskStride,
                                                      sizeof(int32_t) * mv.size(), hipMemcpyDeviceToHost), "D2H");
        if (rc) return rc;
        if (sm[6]) return oops::set_err(ctx, OOPS_ERR_CAPACITY, "device term builder: table overflow");
    }
    for (int q = 0; q < 6; ++q) counts[q] = sm[q];
    if ((int)lv.size() > max_loc || (int)ov.size() > max_ent) return OOPS_ERR_CAPACITY;
    for (size_t l = 0; l < lv.size() && locs; ++l) {
        const int32_t f[6] = {lv[l].nb, lv[l].s, lv[l].t, lv[l].R, lv[l].ent, lv[l].rows};
        for (int q = 0; q < 6; ++q) locs[6 * l + q] = f[q];
    }
    if (val) std::memcpy(val, vv.data(), sizeof(double) * vv.size());
    if (off) std::memcpy(off, ov.data(), sizeof(int32_t) * ov.size());
    if (eslot) std::memcpy(eslot, ev.data(), ev.size());
    if (masks) std::memcpy(masks, mv.data(), sizeof(int32_t) * mv.size());
    return OOPS_OK;
}

int oops_bath_states(uint32_t seed, int32_t nspin, int32_t* states) {
    if (nspin < 0 || (nspin > 0 && !states)) return OOPS_ERR_ARG;
    GlibcRand g(seed);
    for (int32_t i = 0; i < nspin; ++i) states[i] = g.next() % 2;
    return OOPS_OK;
}

int oops_cce_single_sample_terms(const double* coords, int32_t nspin, const int32_t* cluster, int32_t k,
                                 const oops_cce_params* P, const int32_t* states, int32_t which, oops_terms** out,
                                 double* psi) {
    if (!coords || nspin <= 0 || !cluster || k <= 0 || !P || !states || !out || (which != 0 && which != 1))
        return OOPS_ERR_ARG;
    *out = nullptr;
    std::vector<int32_t> cl(cluster, cluster + k);
    for (int i = 0; i < k; ++i)
        if (cl[i] < 0 || cl[i] >= nspin || (i && cl[i] <= cl[i - 1])) return OOPS_ERR_ARG;
    if (k > 20) return OOPS_ERR_UNSUPPORTED;
    NV nv = make_nv(P->nv_pos, P->B);
    SingleSampleProblem sp = build_single_sample(coords, nspin, cl, states, *P, nv);
    std::unique_ptr<oops_terms> t(new oops_terms());
    t->o = which == 0 ? sp.H0 : sp.H1;
    if (psi) std::memcpy(psi, sp.psi.data(), sizeof(cplx) * sp.psi.size());
    *out = t.release();
    return OOPS_OK;
}

int oops_cce_single_sample_batch_create(oops_ctx* ctx, const double* coords, const oops_clusters* C,
                                        const oops_cce_params* P, const int32_t* states, int32_t npulse, int32_t nt,
                                        const double* t, int32_t m, int32_t klim, double tol, int32_t precision,
                                        oops_hahn_batch** out) {
    if (!ctx || !coords || !C || !P || !states || !out || nt < 1 || !t || npulse < 1) return OOPS_ERR_ARG;
    const NV nv = make_nv(P->nv_pos, P->B);
    std::vector<const std::vector<int32_t>*> cls;
    for (int o = 0; o < C->max_order; ++o)
        for (const auto& cl : C->idx[o]) cls.push_back(&cl);
    const size_t nc = cls.size();
    std::vector<SingleSampleProblem> probs(nc);
    auto work = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) probs[i] = build_single_sample(coords, C->nspin, *cls[i], states, *P, nv);
    };
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int T = std::max(1, std::min<int>(threads, (int)(nc / 64) + 1));
    if (T == 1) {
        work(0, nc);
    } else {
        std::vector<std::thread> pool;
        for (int q = 0; q < T; ++q) pool.emplace_back(work, nc * q / T, nc * (q + 1) / T);
        for (auto& th : pool) th.join();
    }
    // two chains per cluster: the ket side runs riffle(H1, H0) and the bra
    // side riffle(H0, H1), both forward from psi; L = Re <psi_bra, psi_ket>
    std::vector<oops_term_table> views(2 * nc);
    std::vector<oops::EchoIn> items(nc);
    for (size_t i = 0; i < nc; ++i) {
        views[2 * i] = probs[i].H0.view();
        views[2 * i + 1] = probs[i].H1.view();
        oops::EchoIn& e = items[i];
        e.opA = &views[2 * i + 1];
        e.opC = &views[2 * i];
        e.opB = &views[2 * i];
        e.opD = &views[2 * i + 1];
        e.u0 = e.y0 = reinterpret_cast<const double*>(probs[i].psi.data());
        e.scale = 1.0;
        e.npulse = npulse;
        e.nt = nt;
        e.t = t;
        e.chains = 2;
    }
    return oops::batch_create(ctx, (int32_t)nc, items.data(), true, m, klim, tol, precision, out);
}

}  // extern "C"

namespace oops {

// Inputs of the device term builder for every cluster of C (false when the
// builder does not take them: clusters above four spins, a polarised bath).
bool cce_builder_inputs(const oops_clusters* C, const oops_cce_params* P, std::vector<BuiltGroup>& groups,
                        int64_t& nitems, TermBuilderArgs& consts, std::vector<std::vector<double>>& rho) {
    if (!C || !P || !device_builder_applies(C, P)) return false;
    bool bad = false;
    groups = builder_groups(C, 0, nullptr, nitems, bad);
    NV nv = make_nv(P->nv_pos, P->B);
    builder_consts(*P, nv, consts, rho);
    return true;
}

// Every cluster's sweep operators (lvA, -lvB^dagger and, for CPMG-N, lvB and
// -lvA^dagger) and rho0, built in parallel over clusters.
int cce_build_items(const double* coords, const oops_clusters* C, const oops_cce_params* P, int32_t npulse,
                    int32_t nt, const double* t, int threads, CceItems& out) {
    if (!coords || !C || !P || nt < 1 || !t || npulse < 1) return OOPS_ERR_ARG;
    struct Hold {
        std::vector<ClusterProblem> probs;
        std::vector<oops_term_table> views;
    };
    auto hold = std::make_shared<Hold>();
    const NV nv = make_nv(P->nv_pos, P->B);
    std::vector<const std::vector<int32_t>*> cls;
    for (int o = 0; o < C->max_order; ++o)
        for (const auto& cl : C->idx[o]) cls.push_back(&cl);
    const size_t nc = cls.size();
    hold->probs.resize(nc);
    const int need = npulse > 1 ? (kNeedEcho | kNeedCpmg) : kNeedEcho;
    auto work = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const auto& cl = *cls[i];
            std::vector<SpinH> spins(cl.size());
            for (size_t q = 0; q < cl.size(); ++q) {
                for (int c = 0; c < 3; ++c) spins[q].r[c] = coords[3 * cl[q] + c];
                isotope_of("13C", spins[q].iso);
            }
            hold->probs[i] = build_problem(spins, *P, nv, need);
        }
    };
    const int T = std::max(1, std::min<int>(threads, (int)(nc / 64) + 1));
    if (T == 1) {
        work(0, nc);
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < T; ++k) pool.emplace_back(work, nc * k / T, nc * (k + 1) / T);
        for (auto& th : pool) th.join();
    }
    hold->views.resize(4 * nc);
    out.t.assign(t, t + nt);
    out.items.resize(nc);
    for (size_t i = 0; i < nc; ++i) {
        ClusterProblem& p = hold->probs[i];
        oops_term_table* v = &hold->views[4 * i];
        v[0] = p.A.view();
        v[1] = p.Bd.view();
        EchoIn& e = out.items[i];
        e.opA = &v[0];
        e.opB = &v[1];
        e.opC = e.opD = nullptr;
        if (npulse > 1) {
            v[2] = p.lvB.view();
            v[3] = p.Ad.view();
            e.opC = &v[2];
            e.opD = &v[3];
        }
        e.u0 = e.y0 = reinterpret_cast<const double*>(p.rho.data());
        e.scale = p.D;
        e.npulse = npulse;
        e.nt = nt;
        e.t = nullptr;  // set below, after out.t is final
    }
    for (auto& e : out.items) e.t = out.t.data();
    out.hold = hold;
    return OOPS_OK;
}

}  // namespace oops
