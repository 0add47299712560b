// A GF2 propagation written against the C++ mirror of the reference's
// interface (nedyson.hpp), in the order DecompSimulation runs it
// (src/molNEsim_decomp_impl.h:63-104): G0_from_h0 initial guess, do_boot,
// then do_tstp for every step, then do_energy.
//
//   gf2_run <dir>
// reads   <dir>/params.txt  "nao naux nt ntau k mu beta dt CorrSteps BootMax BootTol prec device"
//         <dir>/Vija.bin    nao*nao*naux doubles, Vija(i,j,Q)
//         <dir>/Viaj.bin    nao*naux*nao doubles, Viaj(i,Q,j)
//         <dir>/h0.bin, fock.bin   nao*nao doubles
//         <dir>/Gmat.bin    (ntau+1) complex nao x nao blocks
// optional <dir>/equilibrium.txt "nelec" and <dir>/h1.bin (nao*nao doubles):
//         the initial state then comes from MatsubaraSolver (the gfmol
//         role: G^M, mu and the Fock matrix), not from Gmat/mu/fock
// writes  <dir>/Gret.bin, Gles.bin, Gtv.bin  (packed, complex128)
//         <dir>/Gmat_out.bin  G^M used (complex128 blocks)
//         <dir>/energy.txt  one line "t eKin ePot N" per step
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "nedyson.hpp"

using namespace nedyson_amd;

template <class T>
static std::vector<T> load(const std::string &path, size_t count) {
  std::vector<T> v(count);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(count * sizeof(T))))
    throw std::runtime_error("cannot read " + path);
  return v;
}

template <class T>
static void save(const std::string &path, const std::vector<T> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
  if (!f) throw std::runtime_error("cannot write " + path);
}

int main(int argc, char **argv) {
  if (argc != 2) {
    std::cerr << "usage: gf2_run <dir>\n";
    return 2;
  }
  const std::string dir = argv[1];
  try {
    int nao, naux, nt, ntau, k, CorrSteps, BootMax, device;
    double mu, beta, dt, BootTol;
    std::string precname;
    {
      std::ifstream f(dir + "/params.txt");
      if (!(f >> nao >> naux >> nt >> ntau >> k >> mu >> beta >> dt >> CorrSteps >> BootMax >> BootTol >> precname >>
            device))
        throw std::runtime_error("bad params.txt");
    }
    const size_t es = (size_t)nao * nao, np = (size_t)ntau + 1;
    auto Vija = load<double>(dir + "/Vija.bin", es * naux);
    auto Viaj = load<double>(dir + "/Viaj.bin", es * naux);
    auto h0 = load<double>(dir + "/h0.bin", es);
    auto fock = load<double>(dir + "/fock.bin", es);
    auto Gmat = load<cplx>(dir + "/Gmat.bin", np * es);

    System sys(nt, ntau, nao, naux, k, device);
    GREEN G = sys.G();
    molGF2SolverDecomp solver(sys, Vija.data(), Viaj.data());
    dyson Dyson(sys);
    double nelec = 0;
    std::ifstream feq(dir + "/equilibrium.txt");
    if (feq >> nelec) {
      auto h1 = load<double>(dir + "/h1.bin", es);
      MatsubaraSolver mat(sys, h1, nelec, beta);
      int it = mat.run(200, 1e-11);
      mu = mat.mu();
      fock = mat.fock();
      std::cout << "matsubara: " << it << " iterations, mu " << mu << std::endl;
    } else {
      G.write(NED_MAT, 0, (long long)np, Gmat.data());
    }
    G.read(NED_MAT, 0, (long long)np, Gmat.data());
    save(dir + "/Gmat_out.bin", Gmat);
    Dyson.set_matsubara(beta);
    Dyson.G0_from_h0(G, mu, fock.data(), beta, dt);

    DecompSimulation sim(sys, h0, fock, mu, beta, dt, CorrSteps, BootMax, BootTol, prec_of(precname));
    int passes = sim.do_boot();
    std::cout << "bootstrap: " << passes << " passes, converged " << sim.bootstrap_converged << std::endl;
    if (!sim.bootstrap_converged) return 1;
    for (int tstp = k + 1; tstp <= nt; tstp++) sim.do_tstp(tstp);

    std::ofstream fe(dir + "/energy.txt");
    fe.precision(17);
    for (int t = 0; t <= nt; t++) {
      double e[3];
      sim.energy(t, e);
      fe << t << " " << e[0] << " " << e[1] << " " << e[2] << "\n";
    }
    const long long ntri = (long long)(nt + 1) * (nt + 2) / 2;
    std::vector<cplx> buf((size_t)ntri * es);
    G.read(NED_RET, 0, ntri, buf.data());
    save(dir + "/Gret.bin", buf);
    G.read(NED_LES, 0, ntri, buf.data());
    save(dir + "/Gles.bin", buf);
    buf.resize((size_t)(nt + 1) * np * es);
    G.read(NED_TV, 0, (long long)(nt + 1) * (long long)np, buf.data());
    save(dir + "/Gtv.bin", buf);

    // the handles refuse objects of another System
    System other(k, ntau, nao, naux, k, device);
    GREEN Gother = other.G();
    bool refused = false;
    try {
      Dyson.Extrapolate(k + 1, Gother);
    } catch (const Error &e) {
      refused = e.status == NED_ERR_ARG;
    }
    if (!refused) {
      std::cerr << "a GREEN of another System was accepted\n";
      return 1;
    }
    std::cout << "gf2_run ok: " << nt << " steps" << std::endl;
    return 0;
  } catch (const Error &e) {
    std::cerr << "nedyson_amd error " << (int)e.status << ": " << e.what() << std::endl;
    return 1;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
