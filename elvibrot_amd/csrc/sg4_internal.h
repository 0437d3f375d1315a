// Internal declarations shared by the set-up (host C++) and the device plan.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sg4hpsi.h"

namespace sg4 {

struct Basis1D {
  int nq = 0, nb = 0;
  std::vector<double> x, w;         // nodes (affinely scaled), weights
  std::vector<double> d0, d1, d2;   // dnRGB (nq x nb)
  std::vector<double> BtW;          // dnRBGwrho (nb x nq)
  std::vector<double> gg1, gg2;     // dnRGG d1, d2 (nq x nq)
};

void gauss_hermite(int n, std::vector<double>& x, std::vector<double>& wexp);
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

struct Setup {
  int D, L, LB, Lmin;
  std::vector<sg4_mode_desc> modes;
  std::vector<std::vector<Basis1D>> b1;  // [k][l]
  std::vector<int> lam;                  // level of 1-D index, [k][i]
  int64_t nterms = 0, nb_ba = 0;
  std::vector<int32_t> levels;           // D x nterms
  std::vector<double> weights;
  std::vector<int64_t> boff, qoff;
  std::vector<int64_t> cnt, prefix;
  std::vector<int32_t> map;
  std::vector<int32_t> nq_tab, nb_tab;   // [k*(L+1) + l]
  std::vector<double> mats;
  std::vector<double> nodes;
  std::vector<double> qweights;

  Setup(int D, int L, int LB, const sg4_mode_desc* modes);
  void build_terms();
  void build_packed();
  void build_map();
  int64_t rank(const int* idx) const;
  void fill_tables(sg4_tables* t) const;
  void term_grid(int64_t t, double* q) const;
};

void set_error(const std::string& msg);

// plan accessors for the propagator translation unit
void* plan_stream(const sg4_plan* p);
int64_t plan_rows(const sg4_plan* p);
int plan_max_nvec(const sg4_plan* p);
// SG4_OK when ptr is coarse-grained device memory on the plan's GPU, else
// SG4_ERR_ARG with a message naming the argument
int device_block(const sg4_plan* p, const void* ptr, const char* name);

}  // namespace sg4

struct sg4_setup {
  sg4::Setup* s;
};
