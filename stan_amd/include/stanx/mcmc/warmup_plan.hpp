// The warm-up schedule of the metric estimate, computed once per run.
//
// Warm-up transitions are numbered 0, 1, ..., num_warmup - 1.  A transition
// whose number lies in [feed_begin, feed_end) adds its draw to the running
// moments; at each number listed in close_at the moments become the new
// inverse metric and restart.  Between the opening buffer and the closing
// buffer the windows double in length, and a window whose successor would
// not fit is stretched to the end of the estimation span.
//
// The schedule is a plain value: the host samplers read it and the device
// phase machine receives a copy in its launch arguments, so both follow one
// schedule by construction.  It reproduces the windows of the reference's
// windowed_adaptation (src/stan/mcmc/windowed_adaptation.hpp:23-110),
// including two of its corner cases:
//   - num_warmup < 20: nothing is estimated;
//   - buffers that do not fit (init_buffer + window + term_buffer >
//     num_warmup): the span becomes 15% / 75% / 10% of the warm-up, draws
//     are fed over it, but no window ever closes, because the reference
//     leaves its window counter at the value its constructor set
//     (0 + 0 - 1 in unsigned arithmetic) instead of restarting it.
#ifndef STANX_MCMC_WARMUP_PLAN_HPP
#define STANX_MCMC_WARMUP_PLAN_HPP

#include <cstdint>
#include <sstream>
#include <string>

#include "../callbacks.hpp"
#include "../philox.hpp"  // SX_HD

namespace stanx {
namespace mcmc {

struct window_plan {
  static constexpr int kMaxCloses = 40;  // windows double: at most 33 in 2^32 iterations
  unsigned feed_begin;
  unsigned feed_end;
  int n_close;
  unsigned close_at[kMaxCloses];  // ascending
};

// Does warm-up transition `it` feed the running moments?
SX_HD inline bool plan_feeds(const window_plan& p, unsigned it) {
  return it >= p.feed_begin && it < p.feed_end;
}

// Does a window close at transition `it`?  `cursor` is the index of the
// next close not yet passed; the caller advances it when this returns true.
SX_HD inline bool plan_closes(const window_plan& p, int cursor, unsigned it) {
  return cursor < p.n_close && p.close_at[cursor] == it;
}

// The closes of a span [first, last] (last = the final transition that
// feeds) opened by a window of `base` transitions: each window after the
// first is twice as long as the one before, and one whose successor would
// run past `last` ends at `last` instead.  Stops at `limit` (no transition
// numbered `limit` or above happens), and where a computed end does not
// lie after the previous one (a zero-length window): the transition count
// has already passed it, so it never closes.
inline void plan_doubling_windows(window_plan& p, unsigned first, unsigned base,
                                  unsigned last, unsigned limit) {
  p.n_close = 0;
  unsigned len = base;
  unsigned end = first + len - 1;  // unsigned, as the reference computes it
  while (end < limit && p.n_close < window_plan::kMaxCloses) {
    if (p.n_close > 0 && end <= p.close_at[p.n_close - 1]) break;
    p.close_at[p.n_close++] = end;
    if (end == last) break;
    len *= 2;
    const unsigned next = end + len;
    if (next == last || next + 2 * len < last + 1) {
      end = next;
    } else {
      end = last;
    }
  }
}

inline window_plan plan_windows(unsigned num_warmup, unsigned init_buffer,
                                unsigned term_buffer, unsigned base_window,
                                const std::string& estimate, callbacks::logger& logger) {
  window_plan p{};
  p.n_close = 0;
  if (num_warmup < 20) {
    logger.info("WARNING: No " + estimate + " estimation is");
    logger.info("         performed for num_warmup < 20");
    logger.info("");
    return p;  // empty span
  }
  if (init_buffer + base_window + term_buffer > num_warmup) {
    const unsigned ib = static_cast<unsigned>(0.15 * num_warmup);
    const unsigned tb = static_cast<unsigned>(0.10 * num_warmup);
    const unsigned w = num_warmup - (ib + tb);
    logger.info("WARNING: There aren't enough warmup iterations to fit the");
    logger.info("         three stages of adaptation as currently configured.");
    logger.info("         Reducing each adaptation stage to 15%/75%/10% of");
    logger.info("         the given number of warmup iterations:");
    std::ostringstream a, b, c;
    a << "           init_buffer = " << ib;
    b << "           adapt_window = " << w;
    c << "           term_buffer = " << tb;
    logger.info(a.str());
    logger.info(b.str());
    logger.info(c.str());
    logger.info("");
    p.feed_begin = ib;
    p.feed_end = num_warmup - tb;
    return p;  // fed, never closed
  }
  p.feed_begin = init_buffer;
  p.feed_end = num_warmup - term_buffer;
  plan_doubling_windows(p, init_buffer, base_window, num_warmup - term_buffer - 1, num_warmup);
  return p;
}

}  // namespace mcmc
}  // namespace stanx
#endif
