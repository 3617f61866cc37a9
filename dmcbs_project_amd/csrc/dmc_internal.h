// Internal declarations shared by the host set-up (dmc_setup.cpp) and the device code
// (dmc_kernels.hip, dmc_capi.hip).  Not part of the C-ABI.
#pragma once
#include <cstdint>
#include <array>
#include <string>
#include <vector>

#include "dmc_gpu.h"

namespace dmc {

constexpr int kMaxQ = 64;        // plane-wave vectors (k != 0) in the half space
constexpr int kMaxDraw = 1024;   // RNG jump-ahead table length

// Host-side description of one system, built from dmc_params.
struct System {
    dmc_params p{};
    int npart = 0, n = 0;        // particles, particles per spin
    double el = 0, eli = 0;      // box side and its inverse
    double eunit = 1, etrial = 0, nfac = 0;
    double sigma = 0, sig2 = 0, hbar = 0, mass = 1, sqrtm = 1;
    double range = 0, scale = 0; // Jastrow/potential table range and 1/dr
    double rangeb = 0, scaleb = 0; // beta table
    double tab_hbar = 0;         // hbar^2/(2 m_r) of the Jastrow equation
    int ntab = 0, ntab0 = 0, nobeta = 0, cutdrift = 0;
    double dcut_abs = 0;
    // k vectors: q = 0 first, then the half-space vectors shell by shell
    int nk = 0;
    std::vector<int> numsh;                 // half-shell counts, shell 0 = {k = 0}
    std::vector<std::array<int, 3>> ik;     // integer vectors (units of 2 pi / L)
    std::vector<int> kshell;                // shell of each vector
    // tables [0..ntab]
    std::vector<double> vt, ut, dut, d2ut, bt, dbt, d2bt;
    // d/db and d/dd of the beta tables (parameter adjoint)
    std::vector<double> bt_b, dbt_b, d2bt_b, bt_d, dbt_d, d2bt_d;
    double b = 0, d = 0;                    // beta parameters
    std::vector<double> vorb;
};

// The first nsh0 shells of half-space integer vectors in the order of kstates3d
// (kshell.f90:20-156): ik (units of 2 pi / L), numsh[nsh0] vectors per shell.
int enumerate_kshells(int nsh0, std::vector<std::array<int, 3>> &ik, std::vector<int> &numsh,
                      std::string &err);
// The octant vectors of n(k) (setupenk, momentum3d.f90:13-79): ix, iy, iz >= 0 with
// |m|^2 <= ntab in the reference's order, grouped in shells of equal |m|^2 (first index and
// count of each shell).
void enk_shells(int ntab, std::vector<std::array<int, 3>> &ik, std::vector<int> &first,
                std::vector<int> &count);
// Builds everything from the parameters; returns a DMC_* status.
int build_system(const dmc_params &p, System &s, std::string &err);
// Rebuilds the beta table for new (b, d) (initbeta, betafun.f90:74-125).
void build_beta(System &s, double b, double d);
// Initial walkers (dmcbcs.f90:237-301): x[nwalk][npart][3], irn[nwalk].
void initial_walkers(const System &s, int64_t seed, int nwalk, std::vector<double> &x,
                     std::vector<int64_t> &irn);
// Jump-ahead coefficients of ran1: s_t = a[t] * s_0 + c[t] (mod 2^48), t = 0..n-1.
void ran1_jump_table(int n, std::vector<uint64_t> &a, std::vector<uint64_t> &c);
void ran2_jump_table(int n, std::vector<uint64_t> &a, std::vector<uint64_t> &c);

}  // namespace dmc
