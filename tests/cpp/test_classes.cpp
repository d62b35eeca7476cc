// Test of include/oops_classes.hpp (the reference's class interfaces over the
// C ABI), compiled against the header and liboops_amd.so.
//
//   test_classes host <RoyCoord.xyz>   Kronecker algebra and cluster classes (no GPU)
//   test_classes gpu <cases.bin>       MatExpVector on the device against scipy
//                                      expm fixtures (written by the Python test)
//
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>

#include "../../include/oops_classes.hpp"

using namespace oops;

static int g_fail = 0;

static void check(bool ok, const std::string& what, double val = 0.0) {
    std::printf("%s %s (%.3e)\n", ok ? "ok  " : "FAIL", what.c_str(), val);
    if (!ok) g_fail = 1;
}

static cx_mat random_mat(std::mt19937_64& rng, size_t n) {
    std::normal_distribution<double> nd;
    cx_mat m(n, n);
    for (size_t k = 0; k < m.n_elem; ++k) m(k) = cx_double(nd(rng), nd(rng));
    return m;
}

static double max_abs(const cx_mat& a, const cx_mat& b) {
    double e = 0.0;
    for (size_t k = 0; k < a.n_elem; ++k) e = std::max(e, std::abs(a(k) - b(k)));
    return e;
}

static cx_mat matmul(const cx_mat& a, const cx_mat& b) {
    cx_mat c(a.n_rows, b.n_cols);
    for (size_t j = 0; j < b.n_cols; ++j)
        for (size_t l = 0; l < a.n_cols; ++l)
            for (size_t i = 0; i < a.n_rows; ++i) c(i, j) += a(i, l) * b(l, j);
    return c;
}

static cx_mat vec_of(const cx_mat& m) { return cx_mat(m.memptr(), m.n_elem, 1); }

static int host_checks(const char* xyz) {
    std::mt19937_64 rng(7);
    // FLAT / SHARP are the superoperators of left / right multiplication:
    // Flat(m) vec(rho) = vec(m rho), Sharp(m) vec(rho) = vec(rho m)
    const cx_mat m = random_mat(rng, 3), rho = random_mat(rng, 3);
    check(max_abs(matmul(Flat(m), vec_of(rho)), vec_of(matmul(m, rho))) < 1e-12, "Flat(m) vec(rho) = vec(m rho)");
    check(max_abs(matmul(Sharp(m), vec_of(rho)), vec_of(matmul(rho, m))) < 1e-12, "Sharp(m) vec(rho) = vec(rho m)");
    cx_mat mc = m;
    for (size_t k = 0; k < mc.n_elem; ++k) mc(k) = std::conj(mc(k));
    cx_mat mh = mc.st();   // m^dagger
    check(max_abs(matmul(CircleC(m), vec_of(rho)), vec_of(matmul(m, rho) - matmul(rho, mh))) < 1e-12,
          "CircleC(m) vec(rho) = vec(m rho - rho m^dagger)");

    // KronProd: coeff (I (x) A (x) B) on dims (2, 2, 3) with A on spin 1, B on spin 2
    const cx_mat A = random_mat(rng, 2), Bm = random_mat(rng, 3);
    KronProd kp(DIM_LIST{2, 2, 3});
    kp.fill(INDICES{1, 2}, 0.75, TERM{A, Bm});
    const cx_mat F = kp.full();
    double e = 0.0;
    for (size_t r = 0; r < 12; ++r)
        for (size_t c = 0; c < 12; ++c) {
            const size_t r0 = r / 6, r1 = (r / 3) % 2, r2 = r % 3, c0 = c / 6, c1 = (c / 3) % 2, c2 = c % 3;
            const cx_double want = 0.75 * (r0 == c0 ? 1.0 : 0.0) * A(r1, c1) * Bm(r2, c2);
            e = std::max(e, std::abs(F(r, c) - want));
        }
    check(e < 1e-14 && kp.getDim() == 12 && kp.getMatNum() == 2, "KronProd::full element by element", e);
    const cx_vec V = kp.vecterize();
    e = 0.0;
    for (size_t q = 0; q < 4; ++q)
        for (size_t a = 0; a < 4; ++a)
            for (size_t b = 0; b < 9; ++b) {
                const cx_double want = 0.75 * (q == 0 || q == 3 ? 1.0 : 0.0) * A(a) * Bm(b);
                e = std::max(e, std::abs(V((q * 4 + a) * 9 + b) - want));
            }
    check(e < 1e-14 && V.n_elem == 144, "KronProd::vecterize = coeff vec(I) (x) vec(A) (x) vec(B)", e);

    // SumKronProd: full is the sum, scale scales every term, Expand on one spin
    KronProd k2(DIM_LIST{2, 2, 3});
    k2.fill(INDICES{0}, -1.5, TERM{A});
    SumKronProd skp(std::vector<KronProd>{kp, k2});
    check(max_abs(skp.full(), kp.full() + k2.full()) < 1e-14 && skp.getKronProdSize() == 2, "SumKronProd::full");
    SumKronProd one(std::vector<KronProd>{});
    KronProd s1(DIM_LIST{3});
    s1.fill(INDICES{0}, 2.0, TERM{m});
    one.append(s1);
    SumKronProd one_s(std::vector<KronProd>{s1});
    check(max_abs(Expand(one_s, SHARP).full(), Sharp(one_s.full())) < 1e-13, "Expand(SHARP) on one spin");
    check(max_abs(Expand(one_s, FLAT).full(), Flat(one_s.full())) < 1e-13, "Expand(FLAT) on one spin");
    SumKronProd sc(std::vector<KronProd>{kp});
    sc.scale(2.0);
    check(max_abs(sc.full(), cx_double(2.0) * kp.full()) < 1e-14, "SumKronProd::scale");

    // clusters of the reference's RoyCoord.xyz: counts of orders 1-4
    int32_t n = 0;
    oops_read_xyz(xyz, 0, nullptr, nullptr, &n);
    std::vector<double> X(3 * (size_t)n);
    if (oops_read_xyz(xyz, n, X.data(), nullptr, &n) != OOPS_OK) {
        check(false, "read RoyCoord.xyz");
        return 1;
    }
    std::vector<std::array<double, 3>> coords(n);
    for (int i = 0; i < n; ++i) coords[i] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
    cSpinCollection col(coords);
    const size_t want8[4] = {548, 1179, 3530, 12662}, want6[4] = {548, 504, 672, 1107};
    for (double cut : {6.0, 8.0}) {
        cDepthFirstPathTracing dfpt(col.getConnectionMatrix(cut), 4);
        cSpinCluster clst(col, &dfpt);
        clst.make();
        bool ok = clst.getMaxOrder() == 4;
        for (int o = 0; o < 4; ++o) ok = ok && clst.getClusterNum(o) == (cut == 8.0 ? want8[o] : want6[o]);
        check(ok, "cSpinCluster counts on RoyCoord, cutoff " + std::to_string(cut));
        // sub-clusters of a 4-spin cluster: its connected proper subsets (brute force)
        const mat conn = col.getConnectionMatrix(cut);
        const cClusterIndex ci = clst.getClusterIndex(3, clst.getClusterNum(3) / 2);
        const uvec idx = ci.getIndex();
        std::set<std::vector<size_t>> want;
        for (int mask = 1; mask < 15; ++mask) {
            std::vector<size_t> sub;
            for (int q = 0; q < 4; ++q)
                if (mask >> q & 1) sub.push_back(idx(q));
            std::vector<int> seen(sub.size(), 0);
            std::vector<size_t> stack{0};
            seen[0] = 1;
            while (!stack.empty()) {
                const size_t v = stack.back();
                stack.pop_back();
                for (size_t u = 0; u < sub.size(); ++u)
                    if (!seen[u] && conn(sub[v], sub[u]) != 0.0) { seen[u] = 1; stack.push_back(u); }
            }
            if (std::count(seen.begin(), seen.end(), 1) == (long)sub.size()) want.insert(sub);
        }
        std::set<std::vector<size_t>> got;
        for (const auto& pos : clst.getSubClusters(3, clst.getClusterNum(3) / 2)) {
            const uvec u = clst.getClusterIndex(pos).getIndex();
            got.insert(std::vector<size_t>(u.memptr(), u.memptr() + u.n_elem));
        }
        check(got == want && ci.getOrder() == 3 && ci.getNum() == 4,
              "getSubClusters = connected proper subsets, cutoff " + std::to_string(cut));
    }
    return g_fail;
}

// cases.bin: int32 ncase; per case int32 nspin, nterm, nt, then int64 spin_dim[nspin],
// per term: double coeff, int32 nbody, int64 idx[nbody], int64 dim[nbody],
// complex mats (column-major, dim^2 each); complex v[n]; double t[nt]; complex w[n*nt].
static int gpu_checks(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { check(false, "open cases"); return 1; }
    auto rd_i32 = [&]() { int32_t v; f.read((char*)&v, 4); return v; };
    auto rd_i64 = [&]() { int64_t v; f.read((char*)&v, 8); return v; };
    auto rd_f64 = [&]() { double v; f.read((char*)&v, 8); return v; };
    auto rd_c = [&]() { double re = rd_f64(), im = rd_f64(); return cx_double(re, im); };
    const int ncase = rd_i32();
    for (int c = 0; c < ncase; ++c) {
        const int nspin = rd_i32(), nterm = rd_i32(), nt = rd_i32();
        DIM_LIST dims;
        size_t n = 1;
        for (int s = 0; s < nspin; ++s) { dims.push_back((size_t)rd_i64()); n *= dims.back(); }
        std::vector<KronProd> terms;
        for (int t = 0; t < nterm; ++t) {
            const double coeff = rd_f64();
            const int nb = rd_i32();
            INDICES idx;
            std::vector<size_t> d;
            for (int b = 0; b < nb; ++b) idx.push_back((size_t)rd_i64());
            for (int b = 0; b < nb; ++b) d.push_back((size_t)rd_i64());
            TERM mats;
            for (int b = 0; b < nb; ++b) {
                cx_mat M(d[b], d[b]);
                for (size_t k = 0; k < M.n_elem; ++k) M(k) = rd_c();
                mats.push_back(M);
            }
            KronProd kp(dims);
            kp.fill(idx, coeff, mats);
            terms.push_back(kp);
        }
        cx_vec v(n);
        for (size_t i = 0; i < n; ++i) v(i) = rd_c();
        vec tl(nt);
        for (int k = 0; k < nt; ++k) tl(k) = rd_f64();
        cx_mat w(n, nt);
        for (size_t k = 0; k < w.n_elem; ++k) w(k) = rd_c();
        if (!f) { check(false, "read cases"); return 1; }
        double wmax = 0.0;
        for (size_t k = 0; k < w.n_elem; ++k) wmax = std::max(wmax, std::abs(w(k)));
        const SumKronProd skp(terms);
        const std::string tag = "case " + std::to_string(c) + " n=" + std::to_string(n);
        if (n <= 30) {
            // the driver needs m < n; the reference's class fixes m = 30 and stops
            bool threw = false;
            try {
                MatExpVector(skp, v, tl, MatExpVector::InexplicitGPU).run();
            } catch (const std::runtime_error&) {
                threw = true;
            }
            check(threw, tag + ": m = 30 not below n is an error");
            continue;
        }
        for (auto method : {MatExpVector::InexplicitGPU, MatExpVector::Inexplicit, MatExpVector::Explicit,
                            MatExpVector::ExplicitSparse}) {
            MatExpVector mev(skp, v, tl, method);
            const cx_mat got = mev.run();
            const double err = max_abs(got, w) / wmax;
            check(err < 1e-10 && max_abs(mev.getResult(), got) == 0.0,
                  tag + " method " + std::to_string((int)method) + " vs scipy expm", err);
        }
        // the dense and sparse constructors with the prefactor -i
        MatExpVector dense(skp.full(), v, cx_double(0.0, -1.0), tl);
        check(max_abs(dense.run(), w) / wmax < 1e-10, tag + " dense constructor", max_abs(dense.getResult(), w) / wmax);
        const cx_mat full = skp.full();
        sp_cx_mat sp(n, n);
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < n; ++i)
                if (full(i, j) != cx_double(0.0)) sp.insert(i, j, full(i, j));
        MatExpVector sparse(sp, v, cx_double(0.0, -1.0), tl);
        check(max_abs(sparse.run(), w) / wmax < 1e-10, tag + " sparse constructor", max_abs(sparse.getResult(), w) / wmax);
    }
    return g_fail;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s host <xyz> | gpu <cases.bin>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    int rc = mode == "host" ? host_checks(argv[2]) : (mode == "gpu" ? gpu_checks(argv[2]) : 2);
    std::printf("%s\n", rc == 0 ? "all ok" : "FAILED");
    return rc;
}
