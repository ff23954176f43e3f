// Host-only driver for the sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer,
// and ThreadSanitizer for the multithreaded reader).  It exercises the library's host code --
// the polyMesh parser (csrc/fcap_polymesh.cpp) and the restart records
// (csrc/fcap_restart.cpp) -- and the CPU oracle (oracle/fcap_oracle.c) on the meshes it reads.
// Test infrastructure only; built and run by tests/test_sanitizers.py with g++.
//
//   san_driver mesh <polyMesh dir>     read with 1 and 8 threads (results must agree), then
//                                      geometry, CSR, Laplacian and the three solvers of the
//                                      oracle on the mesh
//   san_driver reader <polyMesh dir>   the reader alone, 1 and 8 threads
//   san_driver bad <polyMesh dir>      the read must fail with a message, nothing else
//   san_driver restart <scratch dir>   write / probe / read round trip, then truncated files
// Exit status 0 on success; sanitizer reports make the process fail on their own.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fcap.h"
extern "C" {
#include "fcap_oracle.h"
}

namespace {

int fail(const char *what) {
    std::fprintf(stderr, "san_driver: %s\n", what);
    return 1;
}

struct Mesh {
    int nn = 0, nf = 0, nif = 0, n = 0, nfn = 0, npat = 0;
    std::vector<double> xyz;
    std::vector<int> ptr, fn, own, nei;
};

int read_mesh(const char *dir, int threads, Mesh &m, std::string &err) {
    fc_polymesh *h = nullptr;
    char msg[512] = {0};
    const int rc = fc_polymesh_read(dir, threads, &h, msg, sizeof msg);
    if (rc != FC_OK) {
        err = msg;
        return rc;
    }
    fc_polymesh_sizes(h, &m.nn, &m.nf, &m.nif, &m.n, &m.nfn, &m.npat);
    m.xyz.assign(3 * (size_t)m.nn, 0.0);
    m.ptr.assign(m.nf + 1, 0);
    m.fn.assign(m.nfn > 0 ? m.nfn : 1, 0);
    m.own.assign(m.nf > 0 ? m.nf : 1, 0);
    m.nei.assign(m.nif > 0 ? m.nif : 1, 0);
    int r2 = fc_polymesh_arrays(h, m.xyz.data(), m.ptr.data(), m.fn.data(), m.own.data(), m.nei.data());
    for (int p = 0; p < m.npat && r2 == FC_OK; ++p) {
        char bc[64];
        int nfp = 0, start = 0;
        r2 = fc_polymesh_patch(h, p, bc, sizeof bc, &nfp, &start);
    }
    fc_polymesh_free(h);
    return r2;
}

bool same(const Mesh &a, const Mesh &b) {
    return a.nn == b.nn && a.nf == b.nf && a.nif == b.nif && a.n == b.n && a.xyz == b.xyz && a.ptr == b.ptr &&
           a.fn == b.fn && a.own == b.own && a.nei == b.nei;
}

int run_oracle(const Mesh &m) {
    const int n = m.n, nif = m.nif, nf = m.nf;
    std::vector<double> arx(nf), ary(nf), arz(nf), xf(nf), yf(nf), zf(nf), facint(nf);
    std::vector<double> vol(n), xc(n), yc(n), zc(n);
    if (orc_mesh_geometry(m.nn, m.xyz.data(), nf, nif, m.ptr.data(), m.fn.data(), m.own.data(), m.nei.data(), n,
                          arx.data(), ary.data(), arz.data(), xf.data(), yf.data(), zf.data(), vol.data(), xc.data(),
                          yc.data(), zc.data(), facint.data()) != 0)
        return fail("oracle geometry failed");
    const int nnz = n + 2 * nif;
    std::vector<int> ioff(n + 1), ja(nnz), diag(n), ij(nif), ji(nif);
    if (orc_csr_build(n, nif, m.own.data(), m.nei.data(), ioff.data(), ja.data(), diag.data(), ij.data(),
                      ji.data()) != 0)
        return fail("oracle CSR build failed");
    std::vector<double> mu(n, -1.0), phi(n + (nf - nif), 0.0), a(nnz), su(n), al(1), ar(1);
    unsigned s = 12345u;
    for (int i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        su[i] = ((s >> 8) & 0xffff) / 65536.0 - 0.5;
    }
    // every boundary face a Dirichlet wall (p = 0), so the system is definite (the reader's
    // labels are 1-based, as the oracle takes them)
    const int nwal = nf - nif;
    if (orc_laplacian(n, nif, m.own.data(), m.nei.data(), arx.data(), ary.data(), arz.data(), facint.data(),
                      xc.data(), yc.data(), zc.data(), xf.data(), yf.data(), zf.data(), mu.data(), phi.data(), nwal,
                      nif, n, 0, nif, nullptr, nullptr, nullptr, diag.data(), ij.data(), ji.data(), nnz,
                      a.data(), su.data(), al.data(), ar.data()) != 0)
        return fail("oracle Laplacian failed");
    for (int variant = 0; variant < 3; ++variant) {
        std::vector<double> fi(n, 0.0), res(n), hist(200);
        orc_stats st;
        std::memset(&st, 0, sizeof st);
        int rc = variant == 0   ? orc_iccg(0, n, nnz, ioff.data(), ja.data(), diag.data(), a.data(), su.data(),
                                           fi.data(), 0, nullptr, nullptr, nullptr, nullptr, 1e-8, 200, res.data(),
                                           hist.data(), &st)
                 : variant == 1 ? orc_dpcg(0, n, nnz, ioff.data(), ja.data(), diag.data(), a.data(), su.data(),
                                           fi.data(), 0, nullptr, nullptr, nullptr, nullptr, 1e-8, 200, res.data(),
                                           hist.data(), &st)
                                : orc_bicgstab(0, n, nnz, ioff.data(), ja.data(), diag.data(), a.data(),
                                               su.data(), fi.data(), 0, nullptr, nullptr, nullptr, nullptr, 1e-8,
                                               200, res.data(), hist.data(), &st);
        if (rc < 0 || !std::isfinite(st.resl)) return fail("oracle solver failed");
        std::printf("solver %d: %d iterations, rsm %.3e\n", variant, st.iters, st.rsm);
    }
    return 0;
}

int restart_round_trip(const std::string &dir) {
    const std::string path = dir + "/restart_san";
    char err[512];
    std::vector<std::vector<double>> recs = {std::vector<double>(17), std::vector<double>(1000), {}};
    for (size_t k = 0; k < recs.size(); ++k)
        for (size_t i = 0; i < recs[k].size(); ++i) recs[k][i] = 0.5 * (double)i - (double)k;
    std::vector<long long> len;
    std::vector<const double *> ptrs;
    for (auto &r : recs) {
        len.push_back((long long)r.size());
        ptrs.push_back(r.data());
    }
    if (fc_restart_write(path.c_str(), 7, 0.25, 1, -3.5, (int)recs.size(), len.data(), ptrs.data(), err,
                         sizeof err) != FC_OK)
        return fail(err);
    int itime = 0, nrec = 0;
    double t = 0, g = 0;
    std::vector<long long> flen(8);
    if (fc_restart_probe(path.c_str(), 1, &itime, &t, &g, 8, &nrec, flen.data(), err, sizeof err) != FC_OK)
        return fail(err);
    if (itime != 7 || t != 0.25 || g != -3.5 || nrec != (int)recs.size()) return fail("probe disagrees");
    std::vector<std::vector<double>> back(recs.size());
    std::vector<double *> bptr;
    for (size_t k = 0; k < recs.size(); ++k) {
        back[k].assign(recs[k].size(), -1.0);
        bptr.push_back(back[k].data());
    }
    if (fc_restart_read(path.c_str(), 1, nrec, flen.data(), bptr.data(), err, sizeof err) != FC_OK) return fail(err);
    for (size_t k = 0; k < recs.size(); ++k)
        if (back[k] != recs[k]) return fail("read back differs");
    // truncated copies of the file must be refused, never read past their end
    FILE *f = std::fopen(path.c_str(), "rb");
    std::vector<unsigned char> bytes;
    for (int c; (c = std::fgetc(f)) != EOF;) bytes.push_back((unsigned char)c);
    std::fclose(f);
    const std::string cut = dir + "/restart_cut";
    for (size_t keep : {size_t(0), size_t(3), size_t(20), bytes.size() / 2, bytes.size() - 1}) {
        FILE *o = std::fopen(cut.c_str(), "wb");
        if (keep) std::fwrite(bytes.data(), 1, keep, o);
        std::fclose(o);
        const int rp = fc_restart_probe(cut.c_str(), 1, &itime, &t, &g, 8, &nrec, flen.data(), err, sizeof err);
        const int rr = fc_restart_read(cut.c_str(), 1, (int)recs.size(), len.data(), bptr.data(), err, sizeof err);
        if (rr == FC_OK) return fail("a truncated restart file was read");
        (void)rp;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) return fail("usage: san_driver mesh|reader|bad|restart <path>");
    const std::string mode = argv[1];
    if (mode == "restart") return restart_round_trip(argv[2]);
    Mesh a, b;
    std::string err;
    const int ra = read_mesh(argv[2], 1, a, err);
    if (mode == "bad") {
        if (ra == FC_OK) return fail("a damaged mesh was read");
        if (err.empty()) return fail("no message for a damaged mesh");
        std::printf("refused: %s\n", err.c_str());
        return 0;
    }
    if (ra != FC_OK) return fail(err.c_str());
    if (read_mesh(argv[2], 8, b, err) != FC_OK) return fail(err.c_str());
    if (!same(a, b)) return fail("8 threads read a different mesh than 1");
    std::printf("mesh: %d cells, %d faces\n", a.n, a.nf);
    return mode == "mesh" ? run_oracle(a) : 0;
}
