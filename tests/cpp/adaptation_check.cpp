// Host driver for the warm-up adaptation of the C++ sampler: reads commands
// on stdin and prints one result line per command, so that the pytest side
// can compare against the reference's goldens and against its own
// statement of the reference's window counter.
//   plan NW IB TB BW          -> plan FEED_BEGIN FEED_END N C0 C1 ...
//   tune MU DELTA GAMMA KAPPA T0 N GAP AVG STAT
//                             -> tune EPS N GAP AVG     (one dual-averaging step)
//   varzero D N_LEARN NW IB TB BW
//                             -> varzero V0 ... (metric after N_LEARN zero draws)
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../stan_amd/include/stanx/callbacks.hpp"
#include "../../stan_amd/include/stanx/mcmc/adaptation.hpp"

int main() {
  stanx::callbacks::logger logger;
  std::string line;
  while (std::getline(std::cin, line)) {
    std::istringstream in(line);
    std::string cmd;
    in >> cmd;
    if (cmd == "plan") {
      unsigned nw, ib, tb, bw;
      in >> nw >> ib >> tb >> bw;
      const stanx::mcmc::window_plan p =
          stanx::mcmc::plan_windows(nw, ib, tb, bw, "variance", logger);
      std::printf("plan %u %u %d", p.feed_begin, p.feed_end, p.n_close);
      for (int i = 0; i < p.n_close; ++i) std::printf(" %u", p.close_at[i]);
      std::printf("\n");
    } else if (cmd == "tune") {
      double mu, delta, gamma, kappa, t0, n, gap, avg, stat;
      in >> mu >> delta >> gamma >> kappa >> t0 >> n >> gap >> avg >> stat;
      stanx::mcmc::step_size_tuner t;
      t.set_mu(mu);
      t.set_delta(delta);
      t.set_gamma(gamma);
      t.set_kappa(kappa);
      t.set_t0(t0);
      t.set_iterates(n, gap, avg);
      const double eps = t.update(stat);
      std::printf("tune %.17g %.17g %.17g %.17g\n", eps, t.iterations(), t.gap(), t.average());
    } else if (cmd == "varzero") {
      int D, n_learn;
      unsigned nw, ib, tb, bw;
      in >> D >> n_learn >> nw >> ib >> tb >> bw;
      stanx::mcmc::var_adaptation a(D, "variance");
      a.set_window_params(nw, ib, tb, bw, logger);
      std::vector<double> metric(D, 0.0), q(D, 0.0);
      for (int i = 0; i < n_learn; ++i) a.observe(metric, q);
      std::printf("varzero");
      for (double v : metric) std::printf(" %.17g", v);
      std::printf("\n");
    }
    std::fflush(stdout);
  }
  return 0;
}
