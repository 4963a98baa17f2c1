// On-the-fly reblocking of the report series and automatic shift-damping
// optimisation for the FCIQMC report loop (src/blocking.F90, with the
// blocking_t / blocking_in_t state of src/qmc_data.f90:425-460,700-780).
//
// The reference accumulates, report by report, block sums of the projected
// energy's numerator (sum_j H_0j N_j), its denominator (N_0) and the shift
// for block sizes 2^0 .. 2^lg_max, keeps copies at regular start points,
// chooses the start point whose optimal-block error is smallest, and -- with
// auto_shift_damping -- rescales the shift damping every 50 reports until the
// spread of the instantaneous projected energy matches that of the shift.
// The expressions below keep the reference's arithmetic, including the places
// where it mixes default (single precision) reals into double-precision
// expressions: those decide which start point and block size are chosen, and
// so the damping and every later shift.
#ifndef FQMC_BLOCKING_HPP
#define FQMC_BLOCKING_HPP

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "fqmc_host.h"

namespace fqmc {

// blocking{} options (src/qmc_data.f90:425-460 defaults).
struct BlockingIn {
  bool on_the_fly = false;
  int32_t start_save_frequency = -1;
  int32_t start_point_number = -1;
  int32_t start_point = -1;
  double error_limit = 0.0;
  int32_t min_blocks_used = 10;
  int32_t blocks_used = 2147483647;
  bool auto_shift_damping = false;
  double shift_damping_precision = 2.0;
  bool force_shift_damping_opt = false;
};

// qmc_state_t%shift_damping_status values (src/qmc_data.f90:105-117).
enum ShiftDampingStatus : int32_t {
  kSdNoOptimisation = -2,
  kSdOptimised = -1,
  kSdOptimising = 0,
  kSdWait1 = 1,
  kSdWait2 = 2,
};

class Blocking {
 public:
  static constexpr int kNum = 0, kDen = 1, kShift = 2, kProj = 3;  // dt_numerator .. dt_proj_energy
  static constexpr double kEps = 1.0e-12;                           // depsilon (lib/local/const.F90:90)

  // init_blocking (src/blocking.F90:22-77): nreport is the run's report count.
  void init(int64_t nreport, const BlockingIn& in, int32_t& shift_damping_status) {
    // nint(log(real(nreport))/log(2.0)) + 2 in default reals; nint rounds halves away from zero
    lg_max_ = static_cast<int>(std::lround(std::log(static_cast<float>(nreport)) / std::log(2.0f))) + 2;
    if (in.start_save_frequency < 0) {
      save_fq_ = ipow2(lg_max_ - 10);
      if (save_fq_ < 1) save_fq_ = nreport;
    } else {
      save_fq_ = ipow2(in.start_save_frequency);
    }
    n_saved_startpoints_ = in.start_point_number < 0 ? static_cast<int>(nreport / save_fq_) : in.start_point_number;
    precision_ = in.shift_damping_precision;
    reset();
    if ((in.auto_shift_damping && shift_damping_status == kSdNoOptimisation) || in.force_shift_damping_opt)
      shift_damping_status = kSdWait1;
  }

  // do_blocking (src/blocking.F90:574-635), after a report's estimators are
  // final: proj and d0 are the report's averaged sum_j H_0j N_j and N_0 of the
  // first space, shift its shift.  Returns true when check_error asks the run
  // to stop (qs%reblock_done).
  bool report(int64_t ireport, int64_t ncycles, double proj, double d0, double shift, bool vary_shift,
              const BlockingIn& in) {
    if (start_ireport_ == -1 && in.start_point < 0 && vary_shift) {
      start_ireport_ = ireport;
    } else if (in.start_point >= 0) {
      start_ireport_ = std::lround(static_cast<float>(in.start_point) / static_cast<float>(ncycles));  // nint
    }
    bool done = false;
    if (ireport >= start_ireport_ && start_ireport_ >= 0) {
      n_reports_blocked_ = ireport - start_ireport_ + 1;
      collect_data(proj, d0, shift);
      copy_block();
    }
    if (ireport % 50 == 0 && ireport >= start_ireport_ && start_ireport_ >= 0) {
      change_start(start_point_);
      mean_std_cov();
      find_optimal_block();
      written_mean_ = optimal_mean_[kProj];  // the estimate write_blocking reports
      written_std_ = optimal_std_[kProj];
      done = check_error(ireport, in);
    }
    if (n_reports_blocked_ % save_fq_ == 0 && n_reports_blocked_ > 0) err_comparison();
    return done;
  }

  // update_shift_damping (src/blocking.F90:802-923), one rank.  proj / d0:
  // the report's estimators of the first space; on a change the shift of
  // every space restarts from their ratio and the blocking is reset.  Returns
  // the event: 0 none, 1 damping increased (slow approach), 2 damping
  // rescaled, 3 optimised, 4 waiting for data, 5 waiting to settle.
  int update_shift_damping(int64_t ireport, int32_t& status, double& damping, double proj, double d0,
                           double* shift, int nspaces) {
    if (status == kSdNoOptimisation || status == kSdOptimised) return 0;
    if (ireport < start_ireport_ || start_ireport_ < 0 || ireport % 50 != 0) return 0;
    if (status == kSdWait1) {
      status = kSdOptimising;
      return 0;
    }
    if (status == kSdWait2) {
      status = kSdWait1;
      return 0;
    }
    if (status != kSdOptimising) return 0;
    if (n_reports_blocked_ <= 1) return 0;
    double stds[3], means[3];
    instantaneous(stds, means);
    const double sd_proj = fraction_error(means[kNum], means[kDen], 1, stds[kNum], stds[kDen], block_cov_[0]);
    const double inst = means[kNum] / means[kDen];
    const double diff = std::fabs(inst - means[kShift]);
    bool modified = false;
    int event = 0;
    if (std::fabs(stds[kShift]) < kEps || std::fabs(means[kNum]) < kEps || std::fabs(means[kDen]) < kEps) {
      status = kSdOptimising;
      event = 4;
    } else if (std::fabs(diff) > std::fabs(10 * inst) || std::fabs(diff) > std::fabs(10 * means[kShift])) {
      status = kSdOptimising;
      if (n_increased_damping_ >= 1) {
        --n_increased_damping_;
        event = 5;
      } else {
        damping = std::fmin(0.5, std::fmax(0.01, damping * 5.0));
        modified = true;
        ++n_increased_damping_;
        event = 1;
      }
    } else if (sd_proj / stds[kShift] > precision_ || sd_proj / stds[kShift] < 1.0 / precision_) {
      modified = true;
      damping = damping * 1.0 * sd_proj / stds[kShift];
      status = kSdWait1;
      event = 2;
    } else {
      status = kSdOptimised;
      event = 3;
    }
    if (modified) {  // the shift restarts from the projected energy; reset_blocking_info
      const double s = proj / d0;
      for (int k = 0; k < nspaces; ++k) shift[k] = s;
      reset();
    }
    return event;
  }

  // The optimal-block estimate of the projected energy (optimal_mean /
  // optimal_std of dt_proj_energy) as the last blocking report wrote it
  // (write_blocking, every 50 reports), and the start report of the blocked data.
  double proj_energy_mean() const { return written_mean_; }
  double proj_energy_error() const { return written_std_; }
  int64_t start_ireport() const { return start_ireport_; }

 private:
  struct Cell {
    int32_t n_blocks = 0;
    double acc = 0.0, sum = 0.0, sumsq = 0.0;
  };

  static int64_t ipow2(int e) { return e >= 0 ? (int64_t{1} << e) : 0; }  // 2**e for integers (0 for e < 0)

  Cell& rd(int t, int i) { return data_[t * (lg_max_ + 1) + i]; }
  Cell& rd2(int t, int i) { return data2_[t * (lg_max_ + 1) + i]; }
  Cell& save(int t, int i, int k) { return save_[(k * 3 + t) * (lg_max_ + 1) + i]; }
  double& psave(int i, int k) { return product_save_[k * (lg_max_ + 1) + i]; }

  // a default blocking_t with this run's sizes (allocate_blocking / reset_blocking_info)
  void reset() {
    const int nl = lg_max_ + 1, ns = n_saved_startpoints_ + 1;
    data_.assign(3 * nl, Cell{});
    data2_.assign(3 * nl, Cell{});
    product_.assign(nl, 0.0);
    product2_.assign(nl, 0.0);
    mean_.assign(3 * nl, 0.0);
    std_.assign(3 * nl, 0.0);
    block_cov_.assign(nl, 0.0);
    save_.assign(static_cast<size_t>(3) * nl * ns, Cell{});
    product_save_.assign(static_cast<size_t>(nl) * ns, 0.0);
    err_compare_.assign(static_cast<size_t>(3) * ns, 0.0);
    for (int k = 0; k < 4; ++k) optimal_mean_[k] = optimal_std_[k] = optimal_err_[k] = 0.0;
    for (int k = 0; k < 3; ++k) optimal_size_[k] = 1;
    start_ireport_ = -1;
    start_point_ = 0;
    n_saved_ = 1;
    opt_bl_size_ = 0;
    n_increased_damping_ = 0;
    n_reports_blocked_ = 0;
  }

  // collect_data (src/blocking.F90:206-263)
  void collect_data(double proj, double d0, double shift) {
    for (int i = 0; i <= lg_max_; ++i) {
      rd(kNum, i).acc = rd(kNum, i).acc + proj;
      rd(kDen, i).acc = rd(kDen, i).acc + d0;
      rd(kShift, i).acc = rd(kShift, i).acc + shift;
    }
    for (int i = 0; i <= lg_max_; ++i) {
      const int64_t size = int64_t{1} << i;
      if (n_reports_blocked_ % size != 0) continue;
      for (int t = 0; t < 3; ++t) {
        Cell& c = rd(t, i);
        c.n_blocks = static_cast<int32_t>(n_reports_blocked_ / size);
        const double b = c.acc / static_cast<double>(size);
        c.sum = c.sum + b;
        c.sumsq = c.sumsq + b * b;
      }
      product_[i] = product_[i] + (rd(kNum, i).acc / static_cast<double>(size)) *
                                      (rd(kDen, i).acc / static_cast<double>(size));
      for (int t = 0; t < 3; ++t) rd(t, i).acc = 0.0;
    }
  }

  // mean_std_cov (src/blocking.F90:265-303)
  void mean_std_cov() {
    const int nl = lg_max_ + 1;
    for (int i = 0; i <= lg_max_; ++i) {
      const int32_t nb = rd2(kNum, i).n_blocks;
      for (int t = 0; t < 3; ++t)
        mean_[t * nl + i] = nb > 0 ? rd2(t, i).sum / static_cast<double>(rd2(t, i).n_blocks) : 0.0;
      if (nb > 1) {
        for (int t = 0; t < 3; ++t) {
          const double m = mean_[t * nl + i];
          const double n = static_cast<double>(rd2(t, i).n_blocks);
          double v = rd2(t, i).sumsq / n - m * m;
          v = std::sqrt(v / static_cast<double>(rd2(t, i).n_blocks - 1));
          std_[t * nl + i] = v;
        }
        block_cov_[i] = (product2_[i] - (mean_[kNum * nl + i] * mean_[kDen * nl + i] * static_cast<double>(nb))) /
                        static_cast<double>(nb - 1);
      } else {
        for (int t = 0; t < 3; ++t) std_[t * nl + i] = 0.0;
        block_cov_[i] = 0.0;
      }
    }
  }

  // fraction_error (src/blocking.F90:305-334)
  static double fraction_error(double m1, double m2, int n, double s1, double s2, double cov) {
    const double r = m1 / m2;
    const double a = s1 / m1, b = s2 / m2;
    return std::fabs(r * std::sqrt(a * a + b * b - 2 * cov / (static_cast<double>(n) * m1 * m2)));
  }

  // find_optimal_block (src/blocking.F90:336-424): the Wolff/Lee criterion
  // B^3 > 2 n_blocks B (sigma_B / sigma_0)^4, with the reference's single-
  // precision factors (real(B), the 1.0/3.0 exponent, sqrt(real(...))).
  void find_optimal_block() {
    const int nl = lg_max_ + 1;
    const double third = static_cast<double>(1.0f / 3.0f);
    for (int t = 0; t < 3; ++t) {
      for (int j = 0; j <= lg_max_; ++j) {
        const int32_t B = static_cast<int32_t>(int64_t{1} << j);
        const float nbB = static_cast<float>(2 * rd2(t, j).n_blocks) * static_cast<float>(B);
        const double r = std_[t * nl + j] / std_[t * nl + 0];
        const double r2 = r * r;
        const double crit = std::pow(static_cast<double>(nbB) * (r2 * r2), third);
        if (static_cast<double>(B) > crit) {
          optimal_size_[t] = j;
          break;
        }
      }
      const int os = optimal_size_[t];
      if (os == 1) {
        optimal_mean_[t] = 0.0;
        optimal_std_[t] = 0.0;
      } else {
        optimal_std_[t] = std_[t * nl + os];
        optimal_mean_[t] = std::fabs(optimal_std_[t]) < kEps ? 0.0 : mean_[t * nl + os];
      }
      if (std::fabs(optimal_std_[t]) < kEps) {
        optimal_err_[t] = 0.0;
      } else {
        const float d = std::sqrt(static_cast<float>(2 * (rd2(t, os).n_blocks - 1)));
        optimal_err_[t] = optimal_std_[t] / static_cast<double>(d);
      }
    }
    opt_bl_size_ = optimal_size_[kNum] > optimal_size_[kDen] ? optimal_size_[kNum] : optimal_size_[kDen];
    const int ob = opt_bl_size_;
    if (std::fabs(optimal_mean_[kNum]) < kEps || std::fabs(optimal_mean_[kDen]) < kEps) {
      optimal_mean_[kProj] = optimal_std_[kProj] = optimal_err_[kProj] = 0.0;
    } else {
      optimal_mean_[kProj] = mean_[kNum * nl + ob] / mean_[kDen * nl + ob];
      optimal_std_[kProj] = fraction_error(mean_[kNum * nl + ob], mean_[kDen * nl + ob], rd2(kNum, ob).n_blocks,
                                           std_[kNum * nl + ob], std_[kDen * nl + ob], block_cov_[ob]);
      const double a = optimal_err_[kNum] / optimal_std_[kNum], b = optimal_err_[kDen] / optimal_std_[kDen];
      optimal_err_[kProj] = std::sqrt(a * a + b * b) * optimal_std_[kProj];
    }
  }

  // copy_block (src/blocking.F90:426-446)
  void copy_block() {
    if (n_reports_blocked_ % (save_fq_ * n_saved_) == 0 && n_saved_startpoints_ >= n_saved_) {
      for (int t = 0; t < 3; ++t)
        for (int i = 0; i <= lg_max_; ++i) save(t, i, n_saved_) = rd(t, i);
      for (int i = 0; i <= lg_max_; ++i) psave(i, n_saved_) = product_[i];
      ++n_saved_;
    }
  }

  // change_start (src/blocking.F90:448-517): the blocks with the data before
  // start point `restart` taken out.
  void change_start(int restart) {
    data2_ = data_;
    product2_ = product_;
    if (!(n_reports_blocked_ > save_fq_ * restart)) return;
    for (int i = 0; i <= lg_max_; ++i) {
      if (rd2(kNum, i).n_blocks == 1) {
        for (int t = 0; t < 3; ++t) rd2(t, i) = Cell{};
        product2_[i] = 0.0;
      }
    }
    for (int i = 0; i <= lg_max_; ++i) {
      bool sw = true;
      for (int k = 0; k <= n_saved_startpoints_; ++k) {
        if (std::fabs(save(kNum, i, k).acc) < kEps &&
            static_cast<int64_t>(save(kNum, 0, k).n_blocks) >= static_cast<int64_t>(restart) * save_fq_) {
          for (int t = 0; t < 3; ++t) {
            Cell& c = rd2(t, i);
            const Cell& sv = save(t, i, k);
            c.n_blocks = c.n_blocks - sv.n_blocks;
            c.acc = c.acc - sv.acc;
            c.sum = c.sum - sv.sum;
            c.sumsq = c.sumsq - sv.sumsq;
          }
          product2_[i] = product2_[i] - psave(i, k);
          sw = false;
          break;
        }
      }
      if (sw) {
        for (int t = 0; t < 3; ++t) rd2(t, i) = Cell{};
        product2_[i] = 0.0;
      }
    }
  }

  // err_comparison (src/blocking.F90:519-572)
  void err_comparison() {
    const int ns = n_saved_startpoints_ + 1;
    const float lg_fq = std::log(static_cast<float>(save_fq_)) / std::log(2.0f);
    for (int i = 0; i < ns; ++i) {
      if (n_reports_blocked_ > save_fq_ * i) {
        change_start(i);
        mean_std_cov();
        find_optimal_block();
        for (int j = 0; j < 3; ++j) {
          double& ec = err_compare_[static_cast<size_t>(i) * 3 + j];
          if (std::fabs(optimal_std_[j]) < kEps) {
            ec = 0.0;
            continue;
          }
          const int os = optimal_size_[j];
          const float span = static_cast<float>(n_reports_blocked_ - static_cast<int64_t>(i) * save_fq_);
          int32_t blocks = static_cast<int32_t>(span / static_cast<float>(os));  // int(real(..)/size)
          if (!(static_cast<float>(os - 1) < lg_fq)) blocks -= 1;
          const float rt = std::sqrt(static_cast<float>(blocks * os));
          ec = optimal_err_[j] / (optimal_std_[j] * static_cast<double>(rt));
        }
      } else {
        for (int j = 0; j < 3; ++j) err_compare_[static_cast<size_t>(i) * 3 + j] = 0.0;
      }
    }
    int best = -1;
    for (int j = 0; j < 3; ++j) {
      int m = -1;  // minloc over the positive entries, 0-based (-1: none)
      double v = 0.0;
      for (int i = 0; i < ns; ++i) {
        const double x = err_compare_[static_cast<size_t>(i) * 3 + j];
        if (x > 0.0 && (m < 0 || x < v)) {
          m = i;
          v = x;
        }
      }
      if (m > best) best = m;
    }
    start_point_ = best == -1 ? 0 : best;
  }

  // check_error (src/blocking.F90:682-719); error_sum and number_of_blocks
  // keep their values between calls, as the reference's saved locals do.
  bool check_error(int64_t ireport, const BlockingIn& in) {
    if (optimal_std_[kProj] > 0.0) {
      error_sum_ = optimal_err_[kProj] + optimal_std_[kProj];
      number_of_blocks_ = static_cast<int32_t>(
          static_cast<double>(ireport - (static_cast<int64_t>(start_point_) * save_fq_ + start_ireport_)) /
          std::pow(2.0, opt_bl_size_));
    }
    return (error_sum_ < in.error_limit && number_of_blocks_ > in.min_blocks_used) ||
           number_of_blocks_ > in.blocks_used;
  }

  // get_instantaneous_distributions (src/blocking.F90:925-950)
  void instantaneous(double* sd, double* mean) {
    if (rd2(kNum, 0).n_blocks > 1) {
      for (int t = 0; t < 3; ++t) {
        const double n = static_cast<double>(rd2(t, 0).n_blocks);
        mean[t] = rd2(t, 0).sum / n;
        sd[t] = std::sqrt(rd2(t, 0).sumsq / n - mean[t] * mean[t]);
      }
    } else {
      for (int t = 0; t < 3; ++t) mean[t] = sd[t] = 0.0;
    }
  }

  int lg_max_ = 0, n_saved_startpoints_ = 0;
  int64_t save_fq_ = 1, n_reports_blocked_ = 0, start_ireport_ = -1;
  double precision_ = 2.0;
  std::vector<Cell> data_, data2_, save_;
  std::vector<double> product_, product2_, mean_, std_, block_cov_, product_save_, err_compare_;
  double optimal_mean_[4] = {}, optimal_std_[4] = {}, optimal_err_[4] = {};
  int optimal_size_[3] = {1, 1, 1};
  int start_point_ = 0, n_saved_ = 1, opt_bl_size_ = 0, n_increased_damping_ = 0;
  double error_sum_ = 50.0;
  double written_mean_ = 0.0, written_std_ = 0.0;
  int32_t number_of_blocks_ = 0;
};

// The report loop's shift bookkeeping from one report to the next, shared by
// the drivers: update_shift (src/energy_evaluation.F90:659-711) and the
// vary-shift switch of communicated_energy_estimators (:579-657), the start
// values of init_estimators (src/qmc.F90:880-891, with a restart's shift and
// flags), the soft exit end_report_loop takes once reblocking is done
// (src/qmc_common.F90:1236), and do_blocking / update_shift_damping after the
// report is written (src/fciqmc.f90:427-440).
class ShiftControl {
 public:
  // ntot0: the starting total population of each space; start: a restart's
  // state (may be null).
  void init(const fqh_qmc_in& in, int nspaces, const double* ntot0, const fqh_restart* start) {
    S_ = nspaces;
    tau_ = in.tau;
    ncycles_ = in.mc_cycles;
    target_ = in.target_population;
    proje_ = in.vary_shift_from_proje != 0;
    vary_from_ = in.vary_shift_from;
    damping = in.shift_damping;
    status = start ? start->shift_damping_status : kSdNoOptimisation;
    for (int s = 0; s < S_; ++s) {
      shift[s] = start ? start->shift[s] : in.initial_shift;
      vary[s] = start ? start->vary_shift[s] != 0 : false;
      ntot_old[s] = ntot0[s];
      if (in.vary_shift_set) {
        vary[s] = in.vary_shift != 0;
        if (!vary[s]) shift[s] = in.initial_shift;
      }
    }
    bin_.on_the_fly = in.blocking_on_the_fly != 0;
    bin_.auto_shift_damping = in.auto_shift_damping != 0;
    bin_.shift_damping_precision = in.shift_damping_precision;
    bin_.force_shift_damping_opt = in.force_shift_damping_opt != 0;
    bin_.start_save_frequency = in.start_save_frequency;
    bin_.start_point_number = in.start_point_number;
    bin_.start_point = in.start_point;
    bin_.error_limit = in.error_limit;
    bin_.min_blocks_used = in.min_blocks_used;
    bin_.blocks_used = in.blocks_used;
    reblock_done_ = false;
    if (bin_.on_the_fly) bl_.init(in.nreports, bin_, status);
  }

  // update_shift and the vary-shift switch at the end of a report: ntot, proj
  // and d0 are the report's totals and averaged estimators.  Returns true if
  // the loop is to stop after this report (reblocking converged earlier).
  bool end_report(const double* ntot, const double* proj, const double* d0) {
    const double denom = 1.0 * tau_ * ncycles_;
    for (int s = 0; s < S_; ++s) {
      if (!vary[s]) continue;
      if (target_ <= 0.0)
        shift[s] = shift[s] - std::log(ntot[s] / ntot_old[s]) * damping / denom;
      else  // shift_harmonic_forcing is 0
        shift[s] = shift[s] - std::log(ntot[s] / ntot_old[s]) * damping / denom -
                   std::log(ntot[s] / target_) * 0.0 / denom;
    }
    for (int s = 0; s < S_; ++s) ntot_old[s] = ntot[s];
    for (int s = 0; s < S_; ++s) {
      if (!vary[s] && ntot[s] > target_) {
        vary[s] = true;
        shift[s] = proje_ ? proj[s] / d0[s] : vary_from_;
      }
    }
    return reblock_done_;
  }

  // After the report row: on-the-fly blocking and the shift-damping update
  // (ireport counts this run's reports from 1).
  void after_report(int64_t ireport, double proj0, double d00) {
    if (!bin_.on_the_fly) return;
    if (bl_.report(ireport, ncycles_, proj0, d00, shift[0], vary[0], bin_)) reblock_done_ = true;
    if (bin_.auto_shift_damping) bl_.update_shift_damping(ireport, status, damping, proj0, d00, shift, S_);
  }

  bool blocking() const { return bin_.on_the_fly; }
  const Blocking& blocks() const { return bl_; }

  double shift[FQ_MAX_SPACES] = {0, 0};
  bool vary[FQ_MAX_SPACES] = {false, false};
  double ntot_old[FQ_MAX_SPACES] = {0, 0};
  double damping = 0.05;
  int32_t status = kSdNoOptimisation;

 private:
  int S_ = 1, ncycles_ = 1;
  double tau_ = 0, target_ = 0, vary_from_ = 0;
  bool proje_ = false, reblock_done_ = false;
  BlockingIn bin_;
  Blocking bl_;
};

}  // namespace fqmc

#endif  // FQMC_BLOCKING_HPP
