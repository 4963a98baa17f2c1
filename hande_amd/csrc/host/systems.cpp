// Host-side system set-up: the read-only tables the engine consumes.
//
// Hubbard model in momentum space:
//   src/system.f90:418-584 (init_system: box lengths, reciprocal lattice, U/N)
//   src/basis.f90:258-501 (init_model_basis_fns: k-points in the first
//       Brillouin zone, ranking by kinetic energy, alpha/beta interleaving)
//   src/kpoints.f90:9-53 (calc_kinetic, is_reciprocal_lattice_vector)
//   src/momentum_symmetry.f90:17-115 (sym_table, inv_sym, gamma point)
//   src/reference_determinant.f90:46-210 (Aufbau reference, symmetry search)
//   src/hamiltonian_hub_k.f90:91-110 (slater_condon0_hub_k)
// Molecular integrals from an FCIDUMP-format file (RHF):
//   src/read_in.F90:12-860 (header, orbital ranking, CAS freezing, stores)
//   src/read_in.F90:1129-1216 (get_sp_eigv)
//   src/pg_symmetry.f90:10-98 (irrep counts and lists, XOR products)
//   src/molecular_integrals.F90:452-845,1065-1257 (stores and 8-fold index)
//   src/qmc_common.F90:152-260 (find_single_double_prob)
//   src/hamiltonian_molecular.f90:73-139 (slater_condon0_mol)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "fqmc.h"
#include "fqmc_host.h"

namespace {

constexpr double kPi = 3.1415926535897931;
constexpr double kDepsilon = 1.e-12;
constexpr int kHuge = 2147483647;

int64_t tri_ind(int64_t i, int64_t j) { return (i * (i - 1)) / 2 + j; }  // i >= j, 1-based
int64_t tri_reorder(int64_t i, int64_t j) { return i >= j ? tri_ind(i, j) : tri_ind(j, i); }

// Stable insertion ranking with tolerance (lib/local/ranking.f90:81-138):
// rank[i] (0-based) = index of the i-th smallest element.
std::vector<int> insertion_rank(const std::vector<double>& arr, double tol) {
  const int n = static_cast<int>(arr.size());
  std::vector<int> rank(n);
  for (int i = 0; i < n; ++i) rank[i] = i;
  for (int i = 1; i < n; ++i) {
    int j = i - 1;
    const int tmp = rank[i];
    while (j >= 0) {
      if (arr[rank[j]] - arr[tmp] < tol) break;
      rank[j + 1] = rank[j];
      --j;
    }
    rank[j + 1] = tmp;
  }
  return rank;
}

}  // namespace

struct fqh_system {
  fq_system view{};
  int kind = 0;
  int nbasis = 0, nel = 0, nalpha = 0, nbeta = 0, nwords = 0, ms = 0;
  std::vector<uint64_t> f0;
  std::vector<int> occ0;  // reference occupation list in the reference's order
  std::vector<double> sp_eigv;
  std::vector<int> basis_ms;
  std::vector<int> basis_label;  // hub_k: k-point index of the spatial orbital; read_in: irrep
  // Hubbard
  int ndim = 0;
  std::vector<std::vector<int>> kvec;  // per spatial orbital
  std::vector<int> sym_table, inv_sym;
  int nsym = 0, gamma_sym = 0;
  double coulomb_k = 0.0;
  // molecular
  int norb = 0, nsym_pg = 1;
  std::vector<int> orb_sym, nbasis_sym_spin, sym_spin_basis_fns;
  int max_sym_spin = 0;
  double ecore = 0.0;
  std::vector<double> h1, eri;
  double psingle = 0.0, pdouble = 1.0;
  double H00 = 0.0;

  void set_bit(std::vector<uint64_t>& f, int orb) const {
    f[(orb - 1) / 64] |= (uint64_t{1} << ((orb - 1) % 64));
  }
  bool test_bit(const uint64_t* f, int orb) const { return (f[(orb - 1) / 64] >> ((orb - 1) % 64)) & 1; }
  std::vector<int> decode(const uint64_t* f) const {
    std::vector<int> occ;
    for (int i = 1; i <= nbasis; ++i)
      if (test_bit(f, i)) occ.push_back(i);
    return occ;
  }

  // ---- molecular integral accessors (RHF) ---------------------------------
  int spatial(int i) const { return (i + 1) / 2; }
  double h1_get(int i, int j) const { return h1[(spatial(i) - 1) * norb + (spatial(j) - 1)]; }
  // get_one_body_int_mol_real: zero unless symmetry- and spin-allowed.
  double one_body_real(int i, int j) const {
    if (orb_sym[i - 1] != orb_sym[j - 1] || basis_ms[i - 1] != basis_ms[j - 1]) return 0.0;
    return h1_get(i, j);
  }
  void store_one_body(int i, int j, double x) {
    if (orb_sym[i - 1] != orb_sym[j - 1] || basis_ms[i - 1] != basis_ms[j - 1]) return;
    h1[(spatial(i) - 1) * norb + (spatial(j) - 1)] = x;
    h1[(spatial(j) - 1) * norb + (spatial(i) - 1)] = x;
  }
  int64_t eri_index(int i, int j, int a, int b) const {
    int ii = std::max(i, a), aa = std::min(i, a);
    int jj = std::max(j, b), bb = std::min(j, b);
    int64_t ia = tri_ind(spatial(ii), spatial(aa));
    int64_t jb = tri_ind(spatial(jj), spatial(bb));
    return (ia < jb ? tri_ind(jb, ia) : tri_ind(ia, jb)) - 1;
  }
  bool two_body_allowed(int i, int j, int a, int b) const {
    int sym_ij = orb_sym[i - 1] ^ orb_sym[j - 1];
    int sym_ab = orb_sym[a - 1] ^ orb_sym[b - 1];
    return sym_ij == sym_ab && basis_ms[i - 1] == basis_ms[a - 1] && basis_ms[j - 1] == basis_ms[b - 1];
  }
  double two_body_nonzero(int i, int j, int a, int b) const { return eri[eri_index(i, j, a, b)]; }

  double sc0(const uint64_t* f) const {
    std::vector<int> occ = decode(f);
    if (kind == FQ_SYS_HUB_K) {
      double h = static_cast<double>(nalpha * nbeta) * coulomb_k;
      for (int o : occ) h = h + sp_eigv[o - 1];
      return h;
    }
    double h = ecore;
    for (size_t ie = 0; ie < occ.size(); ++ie) {
      int i = occ[ie];
      h = h + h1_get(i, i);
      for (size_t je = ie + 1; je < occ.size(); ++je) {
        int j = occ[je];
        h = h + two_body_nonzero(i, j, i, j);
        if (basis_ms[i - 1] == basis_ms[j - 1]) h = h - two_body_nonzero(i, j, j, i);
      }
    }
    return h;
  }

  void finish_view() {
    view = fq_system{};
    view.system = kind;
    view.nbasis = nbasis;
    view.nel = nel;
    view.nalpha = nalpha;
    view.nbeta = nbeta;
    view.nvirt = nbasis - nel;
    view.nvirt_alpha = nbasis / 2 - nalpha;
    view.nvirt_beta = nbasis / 2 - nbeta;
    view.nwords = nwords;
    view.f0 = f0.data();
    view.H00 = H00;
    view.sp_eigv = sp_eigv.data();
    view.nsym = nsym;
    view.sym_table = sym_table.empty() ? nullptr : sym_table.data();
    view.inv_sym = inv_sym.empty() ? nullptr : inv_sym.data();
    view.coulomb_k = coulomb_k;
    view.norb = norb;
    view.nsym_pg = nsym_pg;
    view.orb_sym = orb_sym.empty() ? nullptr : orb_sym.data();
    view.nbasis_sym_spin = nbasis_sym_spin.empty() ? nullptr : nbasis_sym_spin.data();
    view.max_sym_spin = max_sym_spin;
    view.sym_spin_basis_fns = sym_spin_basis_fns.empty() ? nullptr : sym_spin_basis_fns.data();
    view.ecore = ecore;
    view.h1 = h1.empty() ? nullptr : h1.data();
    view.eri = eri.empty() ? nullptr : eri.data();
    view.n_eri = static_cast<int64_t>(eri.size());
    view.pattempt_single = psingle;
    view.pattempt_double = pdouble;
  }
};

namespace {

// ---------------------------------------------------------------------------
// Hubbard model in momentum space
// ---------------------------------------------------------------------------
int build_hubbard_k(int nel, int ms, int ndim, const int* lat, double U, double t, int sym_mode, int sym,
                    fqh_system* s) {
  if (ndim < 1 || ndim > 3 || nel < 1) return FQ_ERR_ARG;
  s->kind = FQ_SYS_HUB_K;
  s->ndim = ndim;
  // lattice(i,j) = component i of vector j.
  auto L = [&](int i, int j) { return lat[j * ndim + i]; };
  std::vector<double> box(ndim);
  for (int v = 0; v < ndim; ++v) {
    int dot = 0;
    for (int i = 0; i < ndim; ++i) dot += L(i, v) * L(i, v);
    box[v] = std::sqrt(static_cast<double>(dot));
  }
  double prod = 1.0;
  for (int v = 0; v < ndim; ++v) prod *= box[v];
  const int nsites = static_cast<int>(std::lround(prod));
  // rlattice(:,ivec) = lattice(:,ivec)/box_length(ivec)**2
  std::vector<double> rl(ndim * ndim);  // rl[i*ndim + j] = rlattice(i,j)
  for (int v = 0; v < ndim; ++v)
    for (int i = 0; i < ndim; ++i) rl[i * ndim + v] = L(i, v) / (box[v] * box[v]);
  auto kc_of = [&](const int* k, int i) {
    double acc = 0.0;
    for (int j = 0; j < ndim; ++j) acc = acc + static_cast<double>(k[j]) * rl[i * ndim + j];
    return acc;
  };
  s->coulomb_k = U / nsites;
  s->nel = nel;
  s->ms = ms;
  s->nalpha = (nel + ms) / 2;
  s->nbeta = nel - s->nalpha;
  // k-point enumeration (src/basis.f90:371-408).
  int nmax[3] = {0, 0, 0};
  {
    int limits[3][3] = {{0}};
    for (int i = 0; i < ndim; ++i)
      for (int j = 0; j < ndim; ++j)
        if (L(i, j) != 0)
          limits[i][j] = std::abs(static_cast<int>(std::lround(box[i] * box[i] / (2 * L(i, j)))));
    for (int i = 0; i < ndim; ++i) {
      int m = 0;
      for (int r = 0; r < ndim; ++r) m = std::max(m, limits[r][i]);
      nmax[i] = m;
    }
  }
  std::vector<std::vector<int>> kps;
  for (int kk = -nmax[2]; kk <= nmax[2]; ++kk)
    for (int jj = -nmax[1]; jj <= nmax[1]; ++jj)
      for (int ii = -nmax[0]; ii <= nmax[0]; ++ii) {
        int kp[3] = {ii, jj, kk};
        bool in = true;
        for (int d = 0; d < ndim; ++d) {
          double kc = kc_of(kp, d);
          if (!(kc <= 0.5 + kDepsilon && kc > -0.5 + 1.e-5)) in = false;
        }
        if (in) {
          if (static_cast<int>(kps.size()) == nsites) return FQ_ERR_ARG;
          kps.push_back(std::vector<int>(kp, kp + ndim));
        }
      }
  if (static_cast<int>(kps.size()) != nsites) return FQ_ERR_ARG;
  std::vector<double> eps(nsites);
  for (int n = 0; n < nsites; ++n) {
    double acc = 0.0;
    for (int i = 0; i < ndim; ++i) {
      double kc = 0.0;
      for (int j = 0; j < ndim; ++j) kc = kc + (static_cast<double>(kps[n][j]) + 0.0) * rl[i * ndim + j];
      acc = acc + std::cos(2 * kPi * kc);
    }
    eps[n] = -2 * acc * t;
  }
  std::vector<int> rank = insertion_rank(eps, kDepsilon);
  s->nbasis = 2 * nsites;
  s->nwords = (s->nbasis + 63) / 64;
  s->sp_eigv.resize(s->nbasis);
  s->basis_ms.resize(s->nbasis);
  s->basis_label.resize(s->nbasis);
  s->kvec.resize(nsites);
  for (int i = 0; i < nsites; ++i) {
    s->kvec[i] = kps[rank[i]];
    s->sp_eigv[2 * i] = eps[rank[i]];
    s->sp_eigv[2 * i + 1] = eps[rank[i]];
    s->basis_ms[2 * i] = 1;
    s->basis_ms[2 * i + 1] = -1;
    s->basis_label[2 * i] = i + 1;
    s->basis_label[2 * i + 1] = i + 1;
  }
  // Momentum symmetry (src/momentum_symmetry.f90:17-71).
  s->nsym = nsites;
  s->gamma_sym = 0;
  for (int i = 0; i < nsites; ++i) {
    bool zero = true;
    for (int d = 0; d < ndim; ++d) zero = zero && s->kvec[i][d] == 0;
    if (zero) s->gamma_sym = i + 1;
  }
  if (s->gamma_sym == 0) return FQ_ERR_ARG;
  auto is_rlv = [&](const std::vector<int>& k) {
    bool allz = true;
    for (int d = 0; d < ndim; ++d) allz = allz && k[d] == 0;
    if (allz) return true;
    for (int i = 0; i < ndim; ++i) {
      double kc = kc_of(k.data(), i);
      if (!(std::fabs(kc - std::nearbyint(kc)) < kDepsilon)) return false;
    }
    return true;
  };
  s->sym_table.assign(nsites * nsites, 0);
  s->inv_sym.assign(nsites, 0);
  for (int i = 1; i <= nsites; ++i)
    for (int j = i; j <= nsites; ++j) {
      std::vector<int> ksum(ndim);
      for (int d = 0; d < ndim; ++d) ksum[d] = s->kvec[i - 1][d] + s->kvec[j - 1][d];
      for (int k = 1; k <= nsites; ++k) {
        std::vector<int> diff(ndim);
        for (int d = 0; d < ndim; ++d) diff[d] = ksum[d] - s->kvec[k - 1][d];
        if (is_rlv(diff)) {
          s->sym_table[(i - 1) * nsites + (j - 1)] = k;
          s->sym_table[(j - 1) * nsites + (i - 1)] = k;
          if (k == s->gamma_sym) {
            s->inv_sym[i - 1] = j;
            s->inv_sym[j - 1] = i;
          }
          break;
        }
      }
    }
  int target_sym = kHuge;
  if (sym_mode == FQH_SYM_TOT) target_sym = s->gamma_sym;
  if (sym_mode == FQH_SYM_INDEX) target_sym = sym;
  // Reference determinant (src/reference_determinant.f90:46-210).
  std::vector<int> occ(nel);
  for (int i = 0; i < s->nalpha; ++i) occ[i] = 2 * i + 1;
  for (int i = 0; i < s->nbeta; ++i) occ[i + s->nalpha] = 2 * i + 2;
  auto sym_of = [&](const std::vector<int>& o) {
    int isym = s->gamma_sym;
    for (int orb : o) isym = s->sym_table[((orb + 1) / 2 - 1) * nsites + (isym - 1)];
    return isym;
  };
  if (target_sym != kHuge && target_sym >= 1 && target_sym <= nsites && sym_of(occ) != target_sym) {
    std::vector<uint64_t> f(s->nwords, 0);
    for (int o : occ) s->set_bit(f, o);
    double best = 1.0e300;  // huge(0.0_p) stand-in for the comparison below
    bool found = false;
    std::vector<int> cur = occ;
    for (int ic = 0; ic < nel; ++ic) {
      int i = occ[ic];
      for (int iv = 1; iv <= s->nbasis; ++iv) {
        if (!s->test_bit(f.data(), iv) && s->basis_ms[i - 1] == s->basis_ms[iv - 1]) {
          std::vector<int> tmp = occ;
          tmp[ic] = iv;
          if (sym_of(tmp) == target_sym) {
            double e = 0.0;
            for (int o : tmp) e = e + s->sp_eigv[o - 1];
            if (e + kDepsilon < best) {
              cur = tmp;
              best = e;
              found = true;
            }
          }
        }
      }
    }
    for (int ic = 0; ic < nel; ++ic) {
      int i = occ[ic];
      for (int jc = ic + 1; jc < nel; ++jc) {
        int j = occ[jc];
        for (int iv = 1; iv <= s->nbasis; ++iv) {
          if (s->test_bit(f.data(), iv)) continue;
          for (int jv = iv + 1; jv <= s->nbasis; ++jv) {
            if (!s->test_bit(f.data(), jv) &&
                s->basis_ms[i - 1] + s->basis_ms[j - 1] == s->basis_ms[iv - 1] + s->basis_ms[jv - 1]) {
              std::vector<int> tmp = occ;
              tmp[ic] = iv;
              tmp[jc] = jv;
              if (sym_of(tmp) == target_sym) {
                double e = 0.0;
                for (int o : tmp) e = e + s->sp_eigv[o - 1];
                if (e + kDepsilon < best) {
                  cur = tmp;
                  best = e;
                  found = true;
                }
              }
            }
          }
        }
      }
    }
    if (!found) return FQ_ERR_ARG;
    occ = cur;
  }
  s->occ0 = occ;
  s->f0.assign(s->nwords, 0);
  for (int o : occ) s->set_bit(s->f0, o);
  s->H00 = s->sc0(s->f0.data());
  s->psingle = 0.0;
  s->pdouble = 1.0;
  s->finish_view();
  return FQ_OK;
}

// ---------------------------------------------------------------------------
// Molecular integrals (FCIDUMP format, RHF)
// ---------------------------------------------------------------------------
struct DumpLine {
  double x;
  int i, a, j, b;
};

bool parse_dump(const char* path, int& norb, int& nelec, int& ms2, std::vector<int>& orbsym, bool& uhf,
                std::vector<DumpLine>& lines) {
  std::ifstream in(path);
  if (!in) return false;
  std::string header, line;
  bool in_header = true;
  norb = 0;
  nelec = 0;
  ms2 = kHuge;
  uhf = false;
  while (in_header && std::getline(in, line)) {
    header += line + " ";
    std::string up = line;
    for (char& c : up) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    if (up.find("&END") != std::string::npos || up.find('/') != std::string::npos) in_header = false;
  }
  // Tokenise "KEY=v1,v2,..." pairs of the namelist.
  std::string h;
  for (char c : header) h += static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  auto values_of = [&](const std::string& key) {
    std::vector<std::string> vals;
    size_t p = 0;
    while ((p = h.find(key, p)) != std::string::npos) {
      bool boundary = (p == 0 || !(std::isalnum(static_cast<unsigned char>(h[p - 1]))));
      size_t q = p + key.size();
      while (q < h.size() && h[q] == ' ') ++q;
      if (boundary && q < h.size() && h[q] == '=') {
        ++q;
        std::string cur;
        for (; q < h.size(); ++q) {
          char c = h[q];
          if (std::isalpha(static_cast<unsigned char>(c)) && c != 'E' && c != 'D' && c != 'T' && c != 'F') break;
          if (c == '&') break;
          if (c == ',' || c == ' ') {
            if (!cur.empty()) vals.push_back(cur);
            cur.clear();
            // stop when the next token is a new KEY=
            size_t r = q + 1;
            while (r < h.size() && (h[r] == ' ' || h[r] == ',')) ++r;
            size_t s = r;
            while (s < h.size() && std::isalpha(static_cast<unsigned char>(h[s]))) ++s;
            size_t u = s;
            while (u < h.size() && h[u] == ' ') ++u;
            if (s > r && u < h.size() && h[u] == '=') break;
            continue;
          }
          cur += c;
        }
        if (!cur.empty()) vals.push_back(cur);
        return vals;
      }
      p = q;
    }
    return vals;
  };
  auto v = values_of("NORB");
  if (!v.empty()) norb = std::atoi(v[0].c_str());
  v = values_of("NELEC");
  if (!v.empty()) nelec = std::atoi(v[0].c_str());
  v = values_of("MS2");
  if (!v.empty()) ms2 = std::atoi(v[0].c_str());
  v = values_of("ORBSYM");
  orbsym.assign(norb, 0);
  for (int k = 0; k < norb && k < static_cast<int>(v.size()); ++k) orbsym[k] = std::atoi(v[k].c_str());
  v = values_of("UHF");
  if (!v.empty()) uhf = v[0].find('T') != std::string::npos;
  lines.clear();
  while (std::getline(in, line)) {
    for (char& c : line)
      if (c == 'D' || c == 'd') c = 'E';
    std::istringstream ss(line);
    DumpLine d;
    if (!(ss >> d.x >> d.i >> d.a >> d.j >> d.b)) continue;
    lines.push_back(d);
  }
  return norb > 0;
}

int build_read_in(const char* path, int nel_in, int ms_in, int sym_in, int cas_nel, int cas_norb, fqh_system* s) {
  int norb = 0, nelec = 0, ms2 = 0;
  bool uhf = false;
  std::vector<int> orbsym;
  std::vector<DumpLine> lines;
  if (!parse_dump(path, norb, nelec, ms2, orbsym, uhf, lines)) return FQ_ERR_ARG;
  if (uhf) return FQ_ERR_ARG;  // RHF integral stores only
  s->kind = FQ_SYS_READ_IN;
  int nel = nel_in > 0 ? nel_in : nelec;
  int ms = nel_in > 0 ? ms_in : ms2;
  const int rhf_fac = 2;
  int nbasis = 2 * norb;
  // get_sp_eigv (src/read_in.F90:1129-1216)
  std::vector<double> spe(norb + 1, 0.0);
  {
    const int nocc = nel / 2;
    bool found = false;
    std::vector<char> seen_ijij((norb + 1) * (norb + 1), 0), seen_ijji((norb + 1) * (norb + 1), 0);
    for (const DumpLine& d : lines) {
      int i = d.i, a = d.a, j = d.j, b = d.b;
      double x = d.x;
      if (i > 0 && j == 0 && a == 0 && b == 0) {
        found = true;
        spe[i] = x;
      } else if (!found) {
        if (i == j && a == b && i == a && i > 0) {
          if (i <= nocc) spe[i] = spe[i] + x;
        } else if (i == a && j == b && b > 0) {
          if (!seen_ijij[i * (norb + 1) + j]) {
            seen_ijij[i * (norb + 1) + j] = seen_ijij[j * (norb + 1) + i] = 1;
            if (i <= nocc) spe[j] = spe[j] + 2 * x;
            if (j <= nocc) spe[i] = spe[i] + 2 * x;
          }
        } else if (((i == b && j == a) || (i == j && a == b)) && b > 0) {
          if (!seen_ijji[i * (norb + 1) + a]) {
            seen_ijji[i * (norb + 1) + a] = seen_ijji[a * (norb + 1) + i] = 1;
            if (i <= nocc) spe[a] = spe[a] - x;
            if (a <= nocc) spe[i] = spe[i] - x;
          }
        } else if (i == a && j == 0 && b == 0 && i > 0) {
          spe[i] = spe[i] + x;
        }
      }
    }
  }
  std::vector<double> spe1(spe.begin() + 1, spe.end());
  std::vector<int> r0 = insertion_rank(spe1, kDepsilon);
  std::vector<int> sp_eigv_rank(norb + 1, 0);  // 1-based FCIDUMP index of i-th lowest
  for (int i = 1; i <= norb; ++i) sp_eigv_rank[i] = r0[i - 1] + 1;
  std::vector<int> sp_fcidump_rank(norb + 1, 0);
  for (int i = 0; i <= norb; ++i)
    for (int j = 0; j <= norb; ++j)
      if (sp_eigv_rank[j] == i) {
        sp_fcidump_rank[i] = j;
        break;
      }
  int offset = 0;
  if (cas_nel > 0 && cas_norb > 0) {
    offset = nel - cas_nel;
    nbasis = 2 * cas_norb;
    nel = cas_nel;
  }
  const int norb_act = nbasis / 2;
  s->nbasis = nbasis;
  s->nel = nel;
  s->ms = ms;
  s->nalpha = (nel + ms) / 2;
  s->nbeta = nel - s->nalpha;
  s->nwords = (nbasis + 63) / 64;
  s->norb = norb_act;
  s->sp_eigv.resize(nbasis);
  s->basis_ms.resize(nbasis);
  s->orb_sym.resize(nbasis);
  s->basis_label.resize(nbasis);
  for (int i = 1; i <= norb_act; ++i) {
    int rank = sp_eigv_rank[i + offset / rhf_fac];
    int sym = orbsym[rank - 1] - 1;
    s->orb_sym[2 * i - 1] = sym;
    s->orb_sym[2 * i - 2] = sym;
    s->basis_ms[2 * i - 1] = -1;
    s->basis_ms[2 * i - 2] = 1;
    s->sp_eigv[2 * i - 1] = spe[rank];
    s->sp_eigv[2 * i - 2] = spe[rank];
  }
  int minsym = *std::min_element(s->orb_sym.begin(), s->orb_sym.end());
  if (minsym < 0) std::fill(s->orb_sym.begin(), s->orb_sym.end(), 0);
  s->basis_label = s->orb_sym;
  // init_pg_symmetry without Lz (src/pg_symmetry.f90:10-98); the irrep count
  // is rounded up to a power of two with single-precision logs.
  int maxsym_val = *std::max_element(s->orb_sym.begin(), s->orb_sym.end());
  float lg = std::log(static_cast<float>(maxsym_val + 1)) / std::log(2.0f);
  int maxsym = 1 << static_cast<int>(std::ceil(lg));
  s->nsym_pg = maxsym;
  s->nbasis_sym_spin.assign(2 * maxsym, 0);
  std::vector<int> sym_spin_index(nbasis, 0);
  for (int i = 0; i < nbasis; ++i) {
    int ims = (s->basis_ms[i] + 3) / 2 - 1;
    int& c = s->nbasis_sym_spin[ims * maxsym + s->orb_sym[i]];
    c += 1;
    sym_spin_index[i] = c;
  }
  s->max_sym_spin = *std::max_element(s->nbasis_sym_spin.begin(), s->nbasis_sym_spin.end());
  s->sym_spin_basis_fns.assign(2 * maxsym * s->max_sym_spin, 0);
  for (int i = 0; i < nbasis; ++i) {
    int ims = (s->basis_ms[i] + 3) / 2 - 1;
    int g = ims * maxsym + s->orb_sym[i];
    s->sym_spin_basis_fns[g * s->max_sym_spin + sym_spin_index[i] - 1] = i + 1;
  }
  // Integral stores.
  s->h1.assign(static_cast<size_t>(norb_act) * norb_act, 0.0);
  const int64_t npairs = (static_cast<int64_t>(norb_act) * (norb_act + 1)) / 2;
  s->eri.assign(static_cast<size_t>((npairs * (npairs + 1)) / 2), 0.0);
  double Ecore = 0.0;
  const int self_coulomb_fac = 1;  // RHF
  std::vector<char> seen_iha((static_cast<size_t>(nbasis) * (nbasis + 1)) / 2 + 1, 0);
  std::vector<int> seen_ijij((static_cast<size_t>(offset) * (offset + 1)) / 2 + 1, 0);
  const int64_t ntri = (static_cast<int64_t>(nbasis) * (nbasis + 1)) / 2 + 1;
  std::vector<int> seen_iaib(static_cast<size_t>(offset + 1) * ntri, 0);  // core index -offset+1..0
  auto iaib = [&](int core, int64_t t) -> int& { return seen_iaib[(core + offset - 1 + 1) * ntri + t]; };
  for (const DumpLine& d : lines) {
    double x = d.x;
    int i = rhf_fac * sp_fcidump_rank[d.i];
    int j = rhf_fac * sp_fcidump_rank[d.j];
    int a = rhf_fac * sp_fcidump_rank[d.a];
    int b = rhf_fac * sp_fcidump_rank[d.b];
    int ii = i - offset, jj = j - offset, aa = a - offset, bb = b - offset;
    if (std::max(std::max(ii, jj), std::max(aa, bb)) > nbasis) continue;
    if (i == 0 && j == 0 && a == 0 && b == 0) {
      Ecore = Ecore + x;
    } else if (i > 0 && j == 0 && a == 0 && b == 0) {
      // single-particle eigenvalue line
    } else if (j == 0 && b == 0) {
      if (ii < 1 && ii == aa) {
        Ecore = Ecore + x * rhf_fac;
      } else if (ii > 0 && aa > 0) {
        int64_t t = tri_reorder(ii, aa);
        if (!seen_iha[t]) {
          x = x + s->one_body_real(ii, aa);
          s->store_one_body(ii, aa, x);
          seen_iha[t] = 1;
        }
      }
    } else {
      int orbs[4] = {ii, jj, aa, bb};
      int npos = 0;
      for (int o : orbs) npos += o > 0;
      if (npos == 0) {
        if (ii == aa && jj == bb && ii == jj) {
          int64_t t = tri_reorder(i, j);
          if (seen_ijij[t] % 2 == 0) {
            Ecore = Ecore + x * self_coulomb_fac;
            seen_ijij[t] += 1;
          }
        } else if (ii == aa && jj == bb && ii != jj) {
          int64_t t = tri_reorder(i, j);
          if (seen_ijij[t] % 2 == 0) {
            Ecore = Ecore + x * (rhf_fac * rhf_fac);
            seen_ijij[t] += 1;
          }
        } else if ((ii == bb && jj == aa && ii != jj) || (ii == jj && aa == bb && ii != aa)) {
          int64_t t = (ii == jj) ? tri_reorder(i, a) : tri_reorder(i, j);
          if (seen_ijij[t] < 2) {
            Ecore = Ecore - rhf_fac * x;
            seen_ijij[t] += 2;
          }
        }
      } else if (npos == 2) {
        int active[2], core[2], ia = 0, ic = 0;
        for (int o : orbs) {
          if (o > 0)
            active[ia++] = o;
          else
            core[ic++] = o;
        }
        if (core[0] == core[1]) {
          int64_t t = tri_reorder(active[0], active[1]);
          if ((ii == core[0] && aa == core[0]) || (jj == core[0] && bb == core[0])) {
            if (iaib(core[0], t) % 2 == 0) {
              x = x * rhf_fac + s->one_body_real(active[0], active[1]);
              s->store_one_body(active[0], active[1], x);
              iaib(core[0], t) += 1;
            }
          } else {
            bool gamma = (s->orb_sym[active[0] - 1] ^ s->orb_sym[active[1] - 1]) == 0;
            if (iaib(core[0], t) < 2 && gamma) {
              x = s->one_body_real(active[0], active[1]) - x;
              s->store_one_body(active[0], active[1], x);
              iaib(core[0], t) += 2;
            }
          }
        }
      } else if (npos == 4) {
        if (s->two_body_allowed(ii, jj, aa, bb)) s->eri[s->eri_index(ii, jj, aa, bb)] = x;
      }
    }
  }
  s->ecore = Ecore;
  // Reference determinant: Aufbau, then symmetry search if requested.
  std::vector<int> occ(nel);
  for (int i = 0; i < s->nalpha; ++i) occ[i] = 2 * i + 1;
  for (int i = 0; i < s->nbeta; ++i) occ[i + s->nalpha] = 2 * i + 2;
  auto sym_of = [&](const std::vector<int>& o) {
    int isym = 0;
    for (int orb : o) isym ^= s->orb_sym[orb - 1];
    return isym;
  };
  if (sym_in >= 0 && sym_in < maxsym && sym_of(occ) != sym_in) {
    std::vector<uint64_t> f(s->nwords, 0);
    for (int o : occ) s->set_bit(f, o);
    double best = 1.0e300;
    bool found = false;
    std::vector<int> cur = occ;
    for (int ic = 0; ic < nel; ++ic)
      for (int iv = 1; iv <= nbasis; ++iv)
        if (!s->test_bit(f.data(), iv) && s->basis_ms[occ[ic] - 1] == s->basis_ms[iv - 1]) {
          std::vector<int> tmp = occ;
          tmp[ic] = iv;
          if (sym_of(tmp) == sym_in) {
            double e = 0.0;
            for (int o : tmp) e = e + s->sp_eigv[o - 1];
            if (e + kDepsilon < best) {
              cur = tmp;
              best = e;
              found = true;
            }
          }
        }
    for (int ic = 0; ic < nel; ++ic)
      for (int jc = ic + 1; jc < nel; ++jc)
        for (int iv = 1; iv <= nbasis; ++iv) {
          if (s->test_bit(f.data(), iv)) continue;
          for (int jv = iv + 1; jv <= nbasis; ++jv)
            if (!s->test_bit(f.data(), jv) &&
                s->basis_ms[occ[ic] - 1] + s->basis_ms[occ[jc] - 1] == s->basis_ms[iv - 1] + s->basis_ms[jv - 1]) {
              std::vector<int> tmp = occ;
              tmp[ic] = iv;
              tmp[jc] = jv;
              if (sym_of(tmp) == sym_in) {
                double e = 0.0;
                for (int o : tmp) e = e + s->sp_eigv[o - 1];
                if (e + kDepsilon < best) {
                  cur = tmp;
                  best = e;
                  found = true;
                }
              }
            }
        }
    if (!found) return FQ_ERR_ARG;
    occ = cur;
  }
  s->occ0 = occ;
  s->f0.assign(s->nwords, 0);
  for (int o : occ) s->set_bit(s->f0, o);
  s->H00 = s->sc0(s->f0.data());
  // find_single_double_prob (src/qmc_common.F90:152-260)
  {
    std::vector<int> virt = s->nbasis_sym_spin;
    for (int o : occ) virt[((s->basis_ms[o - 1] + 3) / 2 - 1) * maxsym + s->orb_sym[o - 1]] -= 1;
    int64_t nsingles = 0, ndoubles = 0;
    for (int o : occ) nsingles += virt[((s->basis_ms[o - 1] + 3) / 2 - 1) * maxsym + s->orb_sym[o - 1]];
    for (int ie = 0; ie < nel; ++ie) {
      int ims1 = (s->basis_ms[occ[ie] - 1] + 3) / 2 - 1;
      for (int je = ie + 1; je < nel; ++je) {
        int ims2 = (s->basis_ms[occ[je] - 1] + 3) / 2 - 1;
        for (int isyma = 0; isyma < maxsym; ++isyma) {
          int isymb = isyma ^ (s->orb_sym[occ[ie] - 1] ^ s->orb_sym[occ[je] - 1]);
          if (isyma == isymb) {
            if (ims1 == ims2)
              ndoubles += (virt[ims1 * maxsym + isyma] * (virt[ims2 * maxsym + isymb] - 1)) / 2;
            else
              ndoubles += virt[ims1 * maxsym + isyma] * virt[ims2 * maxsym + isymb];
          } else if (isyma < isymb) {
            ndoubles += virt[ims1 * maxsym + isyma] * virt[ims2 * maxsym + isymb];
            if (ims1 != ims2) ndoubles += virt[ims2 * maxsym + isyma] * virt[ims1 * maxsym + isymb];
          }
        }
      }
    }
    s->psingle = static_cast<double>(nsingles) / static_cast<double>(nsingles + ndoubles);
    s->pdouble = static_cast<double>(ndoubles) / static_cast<double>(nsingles + ndoubles);
  }
  s->finish_view();
  return FQ_OK;
}

}  // namespace

extern "C" {

int fqh_hubbard_k(int nel, int ms, int ndim, const int* lattice, double U, double t, int sym_mode, int sym,
                  fqh_system** out) {
  if (!out || !lattice) return FQ_ERR_ARG;
  auto* s = new fqh_system();
  int rc = build_hubbard_k(nel, ms, ndim, lattice, U, t, sym_mode, sym, s);
  if (rc != FQ_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return FQ_OK;
}

int fqh_read_in(const char* path, int nel, int ms, int sym, int cas_nel, int cas_norb, fqh_system** out) {
  if (!out || !path) return FQ_ERR_ARG;
  auto* s = new fqh_system();
  int rc = build_read_in(path, nel, ms, sym, cas_nel, cas_norb, s);
  if (rc != FQ_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return FQ_OK;
}

const fq_system* fqh_view(const fqh_system* s) { return s ? &s->view : nullptr; }

int fqh_ref_occ(const fqh_system* s, int* occ) {
  if (!s || !occ) return FQ_ERR_ARG;
  for (size_t i = 0; i < s->occ0.size(); ++i) occ[i] = s->occ0[i];
  return FQ_OK;
}

int fqh_basis_info(const fqh_system* s, int* ms, int* label, double* sp_eigv) {
  if (!s) return FQ_ERR_ARG;
  for (int i = 0; i < s->nbasis; ++i) {
    if (ms) ms[i] = s->basis_ms[i];
    if (label) label[i] = s->basis_label[i];
    if (sp_eigv) sp_eigv[i] = s->sp_eigv[i];
  }
  return FQ_OK;
}

double fqh_diagonal(const fqh_system* s, const uint64_t* f) { return s->sc0(f); }

void fqh_free(fqh_system* s) { delete s; }

}  // extern "C"
